NAME          SC205
ROWS
 N  MAXIM
 L  ROW00001
 L  ROW00002
 L  ROW00003
 E  ROW00004
 E  ROW00005
 E  ROW00006
 E  ROW00007
 L  ROW00008
 L  ROW00009
 L  ROW00010
 L  ROW00011
 L  ROW00012
 L  ROW00013
 E  ROW00014
 E  ROW00015
 E  ROW00016
 E  ROW00017
 E  ROW00018
 L  ROW00019
 L  ROW00020
 L  ROW00021
 L  ROW00022
 L  ROW00023
 L  ROW00024
 E  ROW00025
 E  ROW00026
 E  ROW00027
 E  ROW00028
 E  ROW00029
 L  ROW00030
 L  ROW00031
 L  ROW00032
 L  ROW00033
 L  ROW00034
 L  ROW00035
 E  ROW00036
 E  ROW00037
 E  ROW00038
 E  ROW00039
 E  ROW00040
 L  ROW00041
 L  ROW00042
 L  ROW00043
 L  ROW00044
 L  ROW00045
 L  ROW00046
 E  ROW00047
 E  ROW00048
 E  ROW00049
 E  ROW00050
 E  ROW00051
 L  ROW00052
 L  ROW00053
 L  ROW00054
 L  ROW00055
 L  ROW00056
 L  ROW00057
 E  ROW00058
 E  ROW00059
 E  ROW00060
 E  ROW00061
 E  ROW00062
 L  ROW00063
 L  ROW00064
 L  ROW00065
 L  ROW00066
 L  ROW00067
 L  ROW00068
 E  ROW00069
 E  ROW00070
 E  ROW00071
 E  ROW00072
 E  ROW00073
 L  ROW00074
 L  ROW00075
 L  ROW00076
 L  ROW00077
 L  ROW00078
 L  ROW00079
 E  ROW00080
 E  ROW00081
 E  ROW00082
 E  ROW00083
 E  ROW00084
 L  ROW00085
 L  ROW00086
 L  ROW00087
 L  ROW00088
 L  ROW00089
 L  ROW00090
 E  ROW00091
 E  ROW00092
 E  ROW00093
 E  ROW00094
 E  ROW00095
 L  ROW00096
 L  ROW00097
 L  ROW00098
 L  ROW00099
 L  ROW00100
 L  ROW00101
 E  ROW00102
 E  ROW00103
 E  ROW00104
 E  ROW00105
 E  ROW00106
 E  ROW00107
 L  ROW00108
 L  ROW00109
 L  ROW00110
 L  ROW00111
 L  ROW00112
 L  ROW00113
 E  ROW00114
 E  ROW00115
 E  ROW00116
 E  ROW00117
 E  ROW00118
 L  ROW00119
 L  ROW00120
 L  ROW00121
 L  ROW00122
 L  ROW00123
 L  ROW00124
 E  ROW00125
 E  ROW00126
 E  ROW00127
 E  ROW00128
 E  ROW00129
 L  ROW00130
 L  ROW00131
 L  ROW00132
 L  ROW00133
 L  ROW00134
 L  ROW00135
 E  ROW00136
 E  ROW00137
 E  ROW00138
 E  ROW00139
 E  ROW00140
 L  ROW00141
 L  ROW00142
 L  ROW00143
 L  ROW00144
 L  ROW00145
 L  ROW00146
 E  ROW00147
 E  ROW00148
 E  ROW00149
 E  ROW00150
 E  ROW00151
 L  ROW00152
 L  ROW00153
 L  ROW00154
 L  ROW00155
 L  ROW00156
 L  ROW00157
 E  ROW00158
 E  ROW00159
 E  ROW00160
 E  ROW00161
 E  ROW00162
 L  ROW00163
 L  ROW00164
 L  ROW00165
 L  ROW00166
 L  ROW00167
 L  ROW00168
 E  ROW00169
 E  ROW00170
 E  ROW00171
 E  ROW00172
 E  ROW00173
 L  ROW00174
 L  ROW00175
 L  ROW00176
 L  ROW00177
 L  ROW00178
 L  ROW00179
 E  ROW00180
 E  ROW00181
 E  ROW00182
 E  ROW00183
 E  ROW00184
 L  ROW00185
 L  ROW00186
 L  ROW00187
 L  ROW00188
 L  ROW00189
 L  ROW00190
 E  ROW00191
 E  ROW00192
 E  ROW00193
 E  ROW00194
 E  ROW00195
 L  ROW00196
 L  ROW00197
 L  ROW00198
 L  ROW00199
 L  ROW00200
 L  ROW00201
 E  ROW00202
 L  ROW00203
 L  ROW00204
 L  ROW00205
COLUMNS
    COL00001  ROW00001            2.   ROW00002            1.
    COL00001  ROW00005           -1.
    COL00002  ROW00001            1.   ROW00002            2.
    COL00002  ROW00006           -1.
    COL00003  ROW00001           1.5   ROW00002           1.5
    COL00003  ROW00007           -1.
    COL00004  MAXIM              -1.   ROW00004            1.
    COL00004  ROW00014           1.1
    COL00005  ROW00004           -1.   ROW00013            1.
    COL00005  ROW00015            1.
    COL00006  ROW00005            1.   ROW00008           -1.
    COL00006  ROW00016           -1.
    COL00007  ROW00006            1.   ROW00009           -1.
    COL00007  ROW00017           -1.
    COL00008  ROW00007            1.   ROW00010           -1.
    COL00008  ROW00018           -1.
    COL00009  ROW00008            1.   ROW00011           -1.
    COL00010  ROW00009            1.   ROW00012           -1.
    COL00011  ROW00010            1.   ROW00013           -1.
    COL00012  ROW00011            2.   ROW00012            1.
    COL00012  ROW00016           -1.
    COL00013  ROW00011            1.   ROW00012            2.
    COL00013  ROW00017           -1.
    COL00014  ROW00011           1.5   ROW00012           1.5
    COL00014  ROW00018           -1.
    COL00015  ROW00014           -1.   ROW00015            1.
    COL00015  ROW00025           1.1
    COL00016  ROW00015           -1.   ROW00024            1.
    COL00016  ROW00026            1.
    COL00017  ROW00016            1.   ROW00019           -1.
    COL00017  ROW00027           -1.
    COL00018  ROW00017            1.   ROW00020           -1.
    COL00018  ROW00028           -1.
    COL00019  ROW00018            1.   ROW00021           -1.
    COL00019  ROW00029           -1.
    COL00020  ROW00019            1.   ROW00022           -1.
    COL00021  ROW00020            1.   ROW00023           -1.
    COL00022  ROW00021            1.   ROW00024           -1.
    COL00023  ROW00022            2.   ROW00023            1.
    COL00023  ROW00027           -1.
    COL00024  ROW00022            1.   ROW00023            2.
    COL00024  ROW00028           -1.
    COL00025  ROW00022           1.5   ROW00023           1.5
    COL00025  ROW00029           -1.
    COL00026  ROW00025           -1.   ROW00026            1.
    COL00026  ROW00036           1.1
    COL00027  ROW00026           -1.   ROW00035            1.
    COL00027  ROW00037            1.
    COL00028  ROW00027            1.   ROW00030           -1.
    COL00028  ROW00038           -1.
    COL00029  ROW00028            1.   ROW00031           -1.
    COL00029  ROW00039           -1.
    COL00030  ROW00029            1.   ROW00032           -1.
    COL00030  ROW00040           -1.
    COL00031  ROW00030            1.   ROW00033           -1.
    COL00032  ROW00031            1.   ROW00034           -1.
    COL00033  ROW00032            1.   ROW00035           -1.
    COL00034  ROW00033            2.   ROW00034            1.
    COL00034  ROW00038           -1.
    COL00035  ROW00033            1.   ROW00034            2.
    COL00035  ROW00039           -1.
    COL00036  ROW00033           1.5   ROW00034           1.5
    COL00036  ROW00040           -1.
    COL00037  ROW00036           -1.   ROW00037            1.
    COL00037  ROW00047           1.1
    COL00038  ROW00037           -1.   ROW00046            1.
    COL00038  ROW00048            1.
    COL00039  ROW00038            1.   ROW00041           -1.
    COL00039  ROW00049           -1.
    COL00040  ROW00039            1.   ROW00042           -1.
    COL00040  ROW00050           -1.
    COL00041  ROW00040            1.   ROW00043           -1.
    COL00041  ROW00051           -1.
    COL00042  ROW00041            1.   ROW00044           -1.
    COL00043  ROW00042            1.   ROW00045           -1.
    COL00044  ROW00043            1.   ROW00046           -1.
    COL00045  ROW00044            2.   ROW00045            1.
    COL00045  ROW00049           -1.
    COL00046  ROW00044            1.   ROW00045            2.
    COL00046  ROW00050           -1.
    COL00047  ROW00044           1.5   ROW00045           1.5
    COL00047  ROW00051           -1.
    COL00048  ROW00047           -1.   ROW00048            1.
    COL00048  ROW00058           1.1
    COL00049  ROW00048           -1.   ROW00057            1.
    COL00049  ROW00059            1.
    COL00050  ROW00049            1.   ROW00052           -1.
    COL00050  ROW00060           -1.
    COL00051  ROW00050            1.   ROW00053           -1.
    COL00051  ROW00061           -1.
    COL00052  ROW00051            1.   ROW00054           -1.
    COL00052  ROW00062           -1.
    COL00053  ROW00052            1.   ROW00055           -1.
    COL00054  ROW00053            1.   ROW00056           -1.
    COL00055  ROW00054            1.   ROW00057           -1.
    COL00056  ROW00055            2.   ROW00056            1.
    COL00056  ROW00060           -1.
    COL00057  ROW00055            1.   ROW00056            2.
    COL00057  ROW00061           -1.
    COL00058  ROW00055           1.5   ROW00056           1.5
    COL00058  ROW00062           -1.
    COL00059  ROW00058           -1.   ROW00059            1.
    COL00059  ROW00069           1.1
    COL00060  ROW00059           -1.   ROW00068            1.
    COL00060  ROW00070            1.
    COL00061  ROW00060            1.   ROW00063           -1.
    COL00061  ROW00071           -1.
    COL00062  ROW00061            1.   ROW00064           -1.
    COL00062  ROW00072           -1.
    COL00063  ROW00062            1.   ROW00065           -1.
    COL00063  ROW00073           -1.
    COL00064  ROW00063            1.   ROW00066           -1.
    COL00065  ROW00064            1.   ROW00067           -1.
    COL00066  ROW00065            1.   ROW00068           -1.
    COL00067  ROW00066            2.   ROW00067            1.
    COL00067  ROW00071           -1.
    COL00068  ROW00066            1.   ROW00067            2.
    COL00068  ROW00072           -1.
    COL00069  ROW00066           1.5   ROW00067           1.5
    COL00069  ROW00073           -1.
    COL00070  ROW00069           -1.   ROW00070            1.
    COL00070  ROW00080           1.1
    COL00071  ROW00070           -1.   ROW00079            1.
    COL00071  ROW00081            1.
    COL00072  ROW00071            1.   ROW00074           -1.
    COL00072  ROW00082           -1.
    COL00073  ROW00072            1.   ROW00075           -1.
    COL00073  ROW00083           -1.
    COL00074  ROW00073            1.   ROW00076           -1.
    COL00074  ROW00084           -1.
    COL00075  ROW00074            1.   ROW00077           -1.
    COL00076  ROW00075            1.   ROW00078           -1.
    COL00077  ROW00076            1.   ROW00079           -1.
    COL00078  ROW00077            2.   ROW00078            1.
    COL00078  ROW00082           -1.
    COL00079  ROW00077            1.   ROW00078            2.
    COL00079  ROW00083           -1.
    COL00080  ROW00077           1.5   ROW00078           1.5
    COL00080  ROW00084           -1.
    COL00081  ROW00080           -1.   ROW00081            1.
    COL00081  ROW00091           1.1
    COL00082  ROW00081           -1.   ROW00090            1.
    COL00082  ROW00092            1.
    COL00083  ROW00082            1.   ROW00085           -1.
    COL00083  ROW00093           -1.
    COL00084  ROW00083            1.   ROW00086           -1.
    COL00084  ROW00094           -1.
    COL00085  ROW00084            1.   ROW00087           -1.
    COL00085  ROW00095           -1.
    COL00086  ROW00085            1.   ROW00088           -1.
    COL00087  ROW00086            1.   ROW00089           -1.
    COL00088  ROW00087            1.   ROW00090           -1.
    COL00089  ROW00088            2.   ROW00089            1.
    COL00089  ROW00093           -1.
    COL00090  ROW00088            1.   ROW00089            2.
    COL00090  ROW00094           -1.
    COL00091  ROW00088           1.5   ROW00089           1.5
    COL00091  ROW00095           -1.
    COL00092  ROW00091           -1.   ROW00092            1.
    COL00092  ROW00103           1.1
    COL00093  ROW00092           -1.   ROW00101            1.
    COL00093  ROW00104            1.
    COL00094  ROW00093            1.   ROW00096           -1.
    COL00094  ROW00105           -1.
    COL00095  ROW00094            1.   ROW00097           -1.
    COL00095  ROW00106           -1.
    COL00096  ROW00095            1.   ROW00098           -1.
    COL00096  ROW00107           -1.
    COL00097  ROW00096            1.   ROW00099           -1.
    COL00098  ROW00097            1.   ROW00100           -1.
    COL00099  ROW00098            1.   ROW00101           -1.
    COL00100  ROW00099            2.   ROW00100            1.
    COL00100  ROW00105           -1.
    COL00101  ROW00099            1.   ROW00100            2.
    COL00101  ROW00106           -1.
    COL00102  ROW00099           1.5   ROW00100           1.5
    COL00102  ROW00107           -1.
    COL00103  ROW00102            1.
    COL00104  ROW00103           -1.   ROW00104            1.
    COL00104  ROW00114           1.1
    COL00105  ROW00104           -1.   ROW00113            1.
    COL00105  ROW00115            1.
    COL00106  ROW00105            1.   ROW00108           -1.
    COL00106  ROW00116           -1.
    COL00107  ROW00106            1.   ROW00109           -1.
    COL00107  ROW00117           -1.
    COL00108  ROW00107            1.   ROW00110           -1.
    COL00108  ROW00118           -1.
    COL00109  ROW00108            1.   ROW00111           -1.
    COL00110  ROW00109            1.   ROW00112           -1.
    COL00111  ROW00110            1.   ROW00113           -1.
    COL00112  ROW00111            2.   ROW00112            1.
    COL00112  ROW00116           -1.
    COL00113  ROW00111            1.   ROW00112            2.
    COL00113  ROW00117           -1.
    COL00114  ROW00111           1.5   ROW00112           1.5
    COL00114  ROW00118           -1.
    COL00115  ROW00114           -1.   ROW00115            1.
    COL00115  ROW00125           1.1
    COL00116  ROW00115           -1.   ROW00124            1.
    COL00116  ROW00126            1.
    COL00117  ROW00116            1.   ROW00119           -1.
    COL00117  ROW00127           -1.
    COL00118  ROW00117            1.   ROW00120           -1.
    COL00118  ROW00128           -1.
    COL00119  ROW00118            1.   ROW00121           -1.
    COL00119  ROW00129           -1.
    COL00120  ROW00119            1.   ROW00122           -1.
    COL00121  ROW00120            1.   ROW00123           -1.
    COL00122  ROW00121            1.   ROW00124           -1.
    COL00123  ROW00122            2.   ROW00123            1.
    COL00123  ROW00127           -1.
    COL00124  ROW00122            1.   ROW00123            2.
    COL00124  ROW00128           -1.
    COL00125  ROW00122           1.5   ROW00123           1.5
    COL00125  ROW00129           -1.
    COL00126  ROW00125           -1.   ROW00126            1.
    COL00126  ROW00136           1.1
    COL00127  ROW00126           -1.   ROW00135            1.
    COL00127  ROW00137            1.
    COL00128  ROW00127            1.   ROW00130           -1.
    COL00128  ROW00138           -1.
    COL00129  ROW00128            1.   ROW00131           -1.
    COL00129  ROW00139           -1.
    COL00130  ROW00129            1.   ROW00132           -1.
    COL00130  ROW00140           -1.
    COL00131  ROW00130            1.   ROW00133           -1.
    COL00132  ROW00131            1.   ROW00134           -1.
    COL00133  ROW00132            1.   ROW00135           -1.
    COL00134  ROW00133            2.   ROW00134            1.
    COL00134  ROW00138           -1.
    COL00135  ROW00133            1.   ROW00134            2.
    COL00135  ROW00139           -1.
    COL00136  ROW00133           1.5   ROW00134           1.5
    COL00136  ROW00140           -1.
    COL00137  ROW00136           -1.   ROW00137            1.
    COL00137  ROW00147           1.1
    COL00138  ROW00137           -1.   ROW00146            1.
    COL00138  ROW00148            1.
    COL00139  ROW00138            1.   ROW00141           -1.
    COL00139  ROW00149           -1.
    COL00140  ROW00139            1.   ROW00142           -1.
    COL00140  ROW00150           -1.
    COL00141  ROW00140            1.   ROW00143           -1.
    COL00141  ROW00151           -1.
    COL00142  ROW00141            1.   ROW00144           -1.
    COL00143  ROW00142            1.   ROW00145           -1.
    COL00144  ROW00143            1.   ROW00146           -1.
    COL00145  ROW00144            2.   ROW00145            1.
    COL00145  ROW00149           -1.
    COL00146  ROW00144            1.   ROW00145            2.
    COL00146  ROW00150           -1.
    COL00147  ROW00144           1.5   ROW00145           1.5
    COL00147  ROW00151           -1.
    COL00148  ROW00147           -1.   ROW00148            1.
    COL00148  ROW00158           1.1
    COL00149  ROW00148           -1.   ROW00157            1.
    COL00149  ROW00159            1.
    COL00150  ROW00149            1.   ROW00152           -1.
    COL00150  ROW00160           -1.
    COL00151  ROW00150            1.   ROW00153           -1.
    COL00151  ROW00161           -1.
    COL00152  ROW00151            1.   ROW00154           -1.
    COL00152  ROW00162           -1.
    COL00153  ROW00152            1.   ROW00155           -1.
    COL00154  ROW00153            1.   ROW00156           -1.
    COL00155  ROW00154            1.   ROW00157           -1.
    COL00156  ROW00155            2.   ROW00156            1.
    COL00156  ROW00160           -1.
    COL00157  ROW00155            1.   ROW00156            2.
    COL00157  ROW00161           -1.
    COL00158  ROW00155           1.5   ROW00156           1.5
    COL00158  ROW00162           -1.
    COL00159  ROW00158           -1.   ROW00159            1.
    COL00159  ROW00169           1.1
    COL00160  ROW00159           -1.   ROW00168            1.
    COL00160  ROW00170            1.
    COL00161  ROW00160            1.   ROW00163           -1.
    COL00161  ROW00171           -1.
    COL00162  ROW00161            1.   ROW00164           -1.
    COL00162  ROW00172           -1.
    COL00163  ROW00162            1.   ROW00165           -1.
    COL00163  ROW00173           -1.
    COL00164  ROW00163            1.   ROW00166           -1.
    COL00165  ROW00164            1.   ROW00167           -1.
    COL00166  ROW00165            1.   ROW00168           -1.
    COL00167  ROW00166            2.   ROW00167            1.
    COL00167  ROW00171           -1.
    COL00168  ROW00166            1.   ROW00167            2.
    COL00168  ROW00172           -1.
    COL00169  ROW00166           1.5   ROW00167           1.5
    COL00169  ROW00173           -1.
    COL00170  ROW00169           -1.   ROW00170            1.
    COL00170  ROW00180           1.1
    COL00171  ROW00170           -1.   ROW00179            1.
    COL00171  ROW00181            1.
    COL00172  ROW00171            1.   ROW00174           -1.
    COL00172  ROW00182           -1.
    COL00173  ROW00172            1.   ROW00175           -1.
    COL00173  ROW00183           -1.
    COL00174  ROW00173            1.   ROW00176           -1.
    COL00174  ROW00184           -1.
    COL00175  ROW00174            1.   ROW00177           -1.
    COL00176  ROW00175            1.   ROW00178           -1.
    COL00177  ROW00176            1.   ROW00179           -1.
    COL00178  ROW00177            2.   ROW00178            1.
    COL00178  ROW00182           -1.
    COL00179  ROW00177            1.   ROW00178            2.
    COL00179  ROW00183           -1.
    COL00180  ROW00177           1.5   ROW00178           1.5
    COL00180  ROW00184           -1.
    COL00181  ROW00180           -1.   ROW00181            1.
    COL00181  ROW00191           1.1
    COL00182  ROW00181           -1.   ROW00190            1.
    COL00182  ROW00192            1.
    COL00183  ROW00182            1.   ROW00185           -1.
    COL00183  ROW00193           -1.
    COL00184  ROW00183            1.   ROW00186           -1.
    COL00184  ROW00194           -1.
    COL00185  ROW00184            1.   ROW00187           -1.
    COL00185  ROW00195           -1.
    COL00186  ROW00185            1.   ROW00188           -1.
    COL00187  ROW00186            1.   ROW00189           -1.
    COL00188  ROW00187            1.   ROW00190           -1.
    COL00189  ROW00188            2.   ROW00189            1.
    COL00189  ROW00193           -1.
    COL00190  ROW00188            1.   ROW00189            2.
    COL00190  ROW00194           -1.
    COL00191  ROW00188           1.5   ROW00189           1.5
    COL00191  ROW00195           -1.
    COL00192  ROW00191           -1.   ROW00192            1.
    COL00192  ROW00202           1.1
    COL00193  ROW00192           -1.   ROW00201            1.
    COL00194  ROW00193            1.   ROW00196           -1.
    COL00195  ROW00194            1.   ROW00197           -1.
    COL00196  ROW00195            1.   ROW00198           -1.
    COL00197  ROW00196            1.   ROW00199           -1.
    COL00198  ROW00197            1.   ROW00200           -1.
    COL00199  ROW00198            1.   ROW00201           -1.
    COL00200  ROW00199            2.   ROW00200            1.
    COL00200  ROW00203           -.8   ROW00204            .1
    COL00201  ROW00199            1.   ROW00200            2.
    COL00201  ROW00203            .1   ROW00204           -.8
    COL00202  ROW00199           1.5   ROW00200           1.5
    COL00202  ROW00203           .15   ROW00204           .15
    COL00202  ROW00205           -1.
    COL00203  ROW00202           -1.   ROW00205            1.
RHS
    RHS       ROW00001          200.   ROW00002          100.
    RHS       ROW00011          200.   ROW00012          100.
    RHS       ROW00022          200.   ROW00023          100.
    RHS       ROW00033          200.   ROW00034          100.
    RHS       ROW00044          200.   ROW00045          100.
    RHS       ROW00055          200.   ROW00056          100.
    RHS       ROW00066          200.   ROW00067          100.
    RHS       ROW00077          200.   ROW00078          100.
    RHS       ROW00088          200.   ROW00089          100.
    RHS       ROW00099          200.   ROW00100          100.
    RHS       ROW00111          200.   ROW00112          100.
    RHS       ROW00122          200.   ROW00123          100.
    RHS       ROW00133          200.   ROW00134          100.
    RHS       ROW00144          200.   ROW00145          100.
    RHS       ROW00155          200.   ROW00156          100.
    RHS       ROW00166          200.   ROW00167          100.
    RHS       ROW00177          200.   ROW00178          100.
    RHS       ROW00188          200.   ROW00189          100.
    RHS       ROW00199          200.   ROW00200          100.
ENDATA
