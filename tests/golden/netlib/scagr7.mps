NAME          SCAGR7
ROWS
 N  FOB00001
 E  ROW00001
 E  ROW00002
 E  ROW00003
 E  ROW00004
 E  ROW00005
 L  ROW00006
 E  ROW00007
 L  ROW00008
 G  ROW00009
 E  ROW00010
 E  ROW00011
 E  ROW00012
 E  ROW00013
 E  ROW00014
 E  ROW00015
 E  ROW00016
 E  ROW00017
 E  ROW00018
 E  ROW00019
 L  ROW00020
 L  ROW00021
 L  ROW00022
 E  ROW00023
 L  ROW00024
 E  ROW00025
 L  ROW00026
 G  ROW00027
 E  ROW00028
 E  ROW00029
 E  ROW00030
 E  ROW00031
 E  ROW00032
 E  ROW00033
 L  ROW00034
 E  ROW00035
 E  ROW00036
 E  ROW00037
 E  ROW00038
 L  ROW00039
 L  ROW00040
 L  ROW00041
 E  ROW00042
 L  ROW00043
 E  ROW00044
 L  ROW00045
 G  ROW00046
 E  ROW00047
 E  ROW00048
 E  ROW00049
 E  ROW00050
 E  ROW00051
 E  ROW00052
 L  ROW00053
 E  ROW00054
 E  ROW00055
 E  ROW00056
 E  ROW00057
 L  ROW00058
 L  ROW00059
 L  ROW00060
 E  ROW00061
 L  ROW00062
 E  ROW00063
 L  ROW00064
 G  ROW00065
 E  ROW00066
 E  ROW00067
 E  ROW00068
 E  ROW00069
 E  ROW00070
 E  ROW00071
 L  ROW00072
 E  ROW00073
 E  ROW00074
 E  ROW00075
 E  ROW00076
 L  ROW00077
 L  ROW00078
 L  ROW00079
 E  ROW00080
 L  ROW00081
 E  ROW00082
 L  ROW00083
 G  ROW00084
 E  ROW00085
 E  ROW00086
 E  ROW00087
 E  ROW00088
 E  ROW00089
 E  ROW00090
 L  ROW00091
 E  ROW00092
 E  ROW00093
 E  ROW00094
 E  ROW00095
 L  ROW00096
 L  ROW00097
 L  ROW00098
 E  ROW00099
 L  ROW00100
 E  ROW00101
 L  ROW00102
 G  ROW00103
 E  ROW00104
 E  ROW00105
 E  ROW00106
 E  ROW00107
 E  ROW00108
 E  ROW00109
 L  ROW00110
 E  ROW00111
 E  ROW00112
 E  ROW00113
 E  ROW00114
 L  ROW00115
 L  ROW00116
 L  ROW00117
 E  ROW00118
 L  ROW00119
 E  ROW00120
 L  ROW00121
 G  ROW00122
 E  ROW00123
 E  ROW00124
 E  ROW00125
 E  ROW00126
 E  ROW00127
 E  ROW00128
 L  ROW00129
COLUMNS
    COL00001  FOB00001          -35.   ROW00001            1.
    COL00002  FOB00001          54.9   ROW00001            1.
    COL00002  ROW00003            1.   ROW00005           -.4
    COL00002  ROW00015           -.5
    COL00003  FOB00001          54.9   ROW00002            1.
    COL00003  ROW00004            1.   ROW00005           -.4
    COL00003  ROW00015           -.5
    COL00004  FOB00001          -35.   ROW00002            1.
    COL00005  FOB00001          23.5   ROW00003           -1.
    COL00005  ROW00005          -1.7   ROW00010            1.
    COL00005  ROW00013         -.245   ROW00015          -1.4
    COL00006  FOB00001          23.5   ROW00004           -1.
    COL00006  ROW00005          -1.7   ROW00011            1.
    COL00006  ROW00013         -.245   ROW00015          -1.4
    COL00007  FOB00001          8.72   ROW00007            1.
    COL00007  ROW00013            .2
    COL00008  FOB00001          9.72   ROW00005           1.5
    COL00008  ROW00006            1.   ROW00007            1.
    COL00009  FOB00001          6.74   ROW00007            1.
    COL00009  ROW00009            1.   ROW00014           1.2
    COL00010  FOB00001          6.84   ROW00007            1.
    COL00010  ROW00009            1.   ROW00015            1.
    COL00011  FOB00001           15.   ROW00005            1.
    COL00012  FOB00001          22.5   ROW00015            1.
    COL00013  ROW00007            1.   ROW00008            1.
    COL00014  FOB00001         -500.   ROW00010           -1.
    COL00014  ROW00023          -.48   ROW00031          -.49
    COL00015  FOB00001          18.7   ROW00010           -1.
    COL00015  ROW00023          -.48   ROW00030            1.
    COL00015  ROW00031          -.49
    COL00016  FOB00001        -258.3   ROW00011           -1.
    COL00016  ROW00023          -.48   ROW00031          -.49
    COL00017  FOB00001         -662.   ROW00012           -1.
    COL00017  ROW00016           -.5   ROW00017           -.5
    COL00017  ROW00023         -9.32   ROW00030            .7
    COL00017  ROW00031          -.56   ROW00032           -1.
    COL00017  ROW00034            1.
    COL00018  FOB00001            3.   ROW00013           -1.
    COL00018  ROW00020            1.   ROW00031            1.
    COL00019  FOB00001           .39   ROW00014           -1.
    COL00019  ROW00021            1.   ROW00032            1.
    COL00020  FOB00001           .47   ROW00015           -1.
    COL00020  ROW00022            1.   ROW00033            1.
    COL00021  FOB00001          -35.   ROW00016            1.
    COL00022  FOB00001          54.9   ROW00016            1.
    COL00022  ROW00018            1.   ROW00023           -.4
    COL00022  ROW00033           -.5
    COL00023  FOB00001          54.9   ROW00017            1.
    COL00023  ROW00019            1.   ROW00023           -.4
    COL00023  ROW00033           -.5
    COL00024  FOB00001          -35.   ROW00017            1.
    COL00025  FOB00001          23.5   ROW00018           -1.
    COL00025  ROW00023          -1.7   ROW00028            1.
    COL00025  ROW00031         -.245   ROW00033          -1.4
    COL00026  FOB00001          23.5   ROW00019           -1.
    COL00026  ROW00023          -1.7   ROW00029            1.
    COL00026  ROW00031         -.245   ROW00033          -1.4
    COL00027  FOB00001          8.72   ROW00025            1.
    COL00027  ROW00031            .2
    COL00028  FOB00001          9.72   ROW00023           1.5
    COL00028  ROW00024            1.   ROW00025            1.
    COL00029  FOB00001          6.74   ROW00025            1.
    COL00029  ROW00027            1.   ROW00032           1.2
    COL00030  FOB00001          6.84   ROW00025            1.
    COL00030  ROW00027            1.   ROW00033            1.
    COL00031  FOB00001           15.   ROW00023            1.
    COL00032  FOB00001          22.5   ROW00033            1.
    COL00033  ROW00025            1.   ROW00026            1.
    COL00034  FOB00001         -500.   ROW00028           -1.
    COL00034  ROW00042          -.48   ROW00050          -.49
    COL00035  FOB00001          18.7   ROW00028           -1.
    COL00035  ROW00042          -.48   ROW00049            1.
    COL00035  ROW00050          -.49
    COL00036  FOB00001        -258.3   ROW00029           -1.
    COL00036  ROW00042          -.48   ROW00050          -.49
    COL00037  FOB00001         -662.   ROW00030           -1.
    COL00037  ROW00034           -1.   ROW00035           -.5
    COL00037  ROW00036           -.5   ROW00042         -9.32
    COL00037  ROW00049            .7   ROW00050          -.56
    COL00037  ROW00051           -1.   ROW00053            1.
    COL00038  FOB00001            3.   ROW00031           -1.
    COL00038  ROW00039            1.   ROW00050            1.
    COL00039  FOB00001           .39   ROW00032           -1.
    COL00039  ROW00040            1.   ROW00051            1.
    COL00040  FOB00001           .47   ROW00033           -1.
    COL00040  ROW00041            1.   ROW00052            1.
    COL00041  FOB00001          -35.   ROW00035            1.
    COL00042  FOB00001          54.9   ROW00035            1.
    COL00042  ROW00037            1.   ROW00042           -.4
    COL00042  ROW00052           -.5
    COL00043  FOB00001          54.9   ROW00036            1.
    COL00043  ROW00038            1.   ROW00042           -.4
    COL00043  ROW00052           -.5
    COL00044  FOB00001          -35.   ROW00036            1.
    COL00045  FOB00001          23.5   ROW00037           -1.
    COL00045  ROW00042          -1.7   ROW00047            1.
    COL00045  ROW00050         -.245   ROW00052          -1.4
    COL00046  FOB00001          23.5   ROW00038           -1.
    COL00046  ROW00042          -1.7   ROW00048            1.
    COL00046  ROW00050         -.245   ROW00052          -1.4
    COL00047  FOB00001          8.72   ROW00044            1.
    COL00047  ROW00050            .2
    COL00048  FOB00001          9.72   ROW00042           1.5
    COL00048  ROW00043            1.   ROW00044            1.
    COL00049  FOB00001          6.74   ROW00044            1.
    COL00049  ROW00046            1.   ROW00051           1.2
    COL00050  FOB00001          6.84   ROW00044            1.
    COL00050  ROW00046            1.   ROW00052            1.
    COL00051  FOB00001           15.   ROW00042            1.
    COL00052  FOB00001          22.5   ROW00052            1.
    COL00053  ROW00044            1.   ROW00045            1.
    COL00054  FOB00001         -500.   ROW00047           -1.
    COL00054  ROW00061          -.48   ROW00069          -.49
    COL00055  FOB00001          18.7   ROW00047           -1.
    COL00055  ROW00061          -.48   ROW00068            1.
    COL00055  ROW00069          -.49
    COL00056  FOB00001        -258.3   ROW00048           -1.
    COL00056  ROW00061          -.48   ROW00069          -.49
    COL00057  FOB00001         -662.   ROW00049           -1.
    COL00057  ROW00053           -1.   ROW00054           -.5
    COL00057  ROW00055           -.5   ROW00061         -9.32
    COL00057  ROW00068            .7   ROW00069          -.56
    COL00057  ROW00070           -1.   ROW00072            1.
    COL00058  FOB00001            3.   ROW00050           -1.
    COL00058  ROW00058            1.   ROW00069            1.
    COL00059  FOB00001           .39   ROW00051           -1.
    COL00059  ROW00059            1.   ROW00070            1.
    COL00060  FOB00001           .47   ROW00052           -1.
    COL00060  ROW00060            1.   ROW00071            1.
    COL00061  FOB00001          -35.   ROW00054            1.
    COL00062  FOB00001          54.9   ROW00054            1.
    COL00062  ROW00056            1.   ROW00061           -.4
    COL00062  ROW00071           -.5
    COL00063  FOB00001          54.9   ROW00055            1.
    COL00063  ROW00057            1.   ROW00061           -.4
    COL00063  ROW00071           -.5
    COL00064  FOB00001          -35.   ROW00055            1.
    COL00065  FOB00001          23.5   ROW00056           -1.
    COL00065  ROW00061          -1.7   ROW00066            1.
    COL00065  ROW00069         -.245   ROW00071          -1.4
    COL00066  FOB00001          23.5   ROW00057           -1.
    COL00066  ROW00061          -1.7   ROW00067            1.
    COL00066  ROW00069         -.245   ROW00071          -1.4
    COL00067  FOB00001          8.72   ROW00063            1.
    COL00067  ROW00069            .2
    COL00068  FOB00001          9.72   ROW00061           1.5
    COL00068  ROW00062            1.   ROW00063            1.
    COL00069  FOB00001          6.74   ROW00063            1.
    COL00069  ROW00065            1.   ROW00070           1.2
    COL00070  FOB00001          6.84   ROW00063            1.
    COL00070  ROW00065            1.   ROW00071            1.
    COL00071  FOB00001           15.   ROW00061            1.
    COL00072  FOB00001          22.5   ROW00071            1.
    COL00073  ROW00063            1.   ROW00064            1.
    COL00074  FOB00001         -500.   ROW00066           -1.
    COL00074  ROW00080          -.48   ROW00088          -.49
    COL00075  FOB00001          18.7   ROW00066           -1.
    COL00075  ROW00080          -.48   ROW00087            1.
    COL00075  ROW00088          -.49
    COL00076  FOB00001        -258.3   ROW00067           -1.
    COL00076  ROW00080          -.48   ROW00088          -.49
    COL00077  FOB00001         -662.   ROW00068           -1.
    COL00077  ROW00072           -1.   ROW00073           -.5
    COL00077  ROW00074           -.5   ROW00080         -9.32
    COL00077  ROW00087            .7   ROW00088          -.56
    COL00077  ROW00089           -1.   ROW00091            1.
    COL00078  FOB00001            3.   ROW00069           -1.
    COL00078  ROW00077            1.   ROW00088            1.
    COL00079  FOB00001           .39   ROW00070           -1.
    COL00079  ROW00078            1.   ROW00089            1.
    COL00080  FOB00001           .47   ROW00071           -1.
    COL00080  ROW00079            1.   ROW00090            1.
    COL00081  FOB00001          -35.   ROW00073            1.
    COL00082  FOB00001          54.9   ROW00073            1.
    COL00082  ROW00075            1.   ROW00080           -.4
    COL00082  ROW00090           -.5
    COL00083  FOB00001          54.9   ROW00074            1.
    COL00083  ROW00076            1.   ROW00080           -.4
    COL00083  ROW00090           -.5
    COL00084  FOB00001          -35.   ROW00074            1.
    COL00085  FOB00001          23.5   ROW00075           -1.
    COL00085  ROW00080          -1.7   ROW00085            1.
    COL00085  ROW00088         -.245   ROW00090          -1.4
    COL00086  FOB00001          23.5   ROW00076           -1.
    COL00086  ROW00080          -1.7   ROW00086            1.
    COL00086  ROW00088         -.245   ROW00090          -1.4
    COL00087  FOB00001          8.72   ROW00082            1.
    COL00087  ROW00088            .2
    COL00088  FOB00001          9.72   ROW00080           1.5
    COL00088  ROW00081            1.   ROW00082            1.
    COL00089  FOB00001          6.74   ROW00082            1.
    COL00089  ROW00084            1.   ROW00089           1.2
    COL00090  FOB00001          6.84   ROW00082            1.
    COL00090  ROW00084            1.   ROW00090            1.
    COL00091  FOB00001           15.   ROW00080            1.
    COL00092  FOB00001          22.5   ROW00090            1.
    COL00093  ROW00082            1.   ROW00083            1.
    COL00094  FOB00001         -500.   ROW00085           -1.
    COL00094  ROW00099          -.48   ROW00107          -.49
    COL00095  FOB00001          18.7   ROW00085           -1.
    COL00095  ROW00099          -.48   ROW00106            1.
    COL00095  ROW00107          -.49
    COL00096  FOB00001        -258.3   ROW00086           -1.
    COL00096  ROW00099          -.48   ROW00107          -.49
    COL00097  FOB00001         -662.   ROW00087           -1.
    COL00097  ROW00091           -1.   ROW00092           -.5
    COL00097  ROW00093           -.5   ROW00099         -9.32
    COL00097  ROW00106            .7   ROW00107          -.56
    COL00097  ROW00108           -1.   ROW00110            1.
    COL00098  FOB00001            3.   ROW00088           -1.
    COL00098  ROW00096            1.   ROW00107            1.
    COL00099  FOB00001           .39   ROW00089           -1.
    COL00099  ROW00097            1.   ROW00108            1.
    COL00100  FOB00001           .47   ROW00090           -1.
    COL00100  ROW00098            1.   ROW00109            1.
    COL00101  FOB00001          -35.   ROW00092            1.
    COL00102  FOB00001          54.9   ROW00092            1.
    COL00102  ROW00094            1.   ROW00099           -.4
    COL00102  ROW00109           -.5
    COL00103  FOB00001          54.9   ROW00093            1.
    COL00103  ROW00095            1.   ROW00099           -.4
    COL00103  ROW00109           -.5
    COL00104  FOB00001          -35.   ROW00093            1.
    COL00105  FOB00001          23.5   ROW00094           -1.
    COL00105  ROW00099          -1.7   ROW00104            1.
    COL00105  ROW00107         -.245   ROW00109          -1.4
    COL00106  FOB00001          23.5   ROW00095           -1.
    COL00106  ROW00099          -1.7   ROW00105            1.
    COL00106  ROW00107         -.245   ROW00109          -1.4
    COL00107  FOB00001          8.72   ROW00101            1.
    COL00107  ROW00107            .2
    COL00108  FOB00001          9.72   ROW00099           1.5
    COL00108  ROW00100            1.   ROW00101            1.
    COL00109  FOB00001          6.74   ROW00101            1.
    COL00109  ROW00103            1.   ROW00108           1.2
    COL00110  FOB00001          6.84   ROW00101            1.
    COL00110  ROW00103            1.   ROW00109            1.
    COL00111  FOB00001           15.   ROW00099            1.
    COL00112  FOB00001          22.5   ROW00109            1.
    COL00113  ROW00101            1.   ROW00102            1.
    COL00114  FOB00001         -500.   ROW00104           -1.
    COL00114  ROW00118          -.48   ROW00126          -.49
    COL00115  FOB00001          18.7   ROW00104           -1.
    COL00115  ROW00118          -.48   ROW00125            1.
    COL00115  ROW00126          -.49
    COL00116  FOB00001        -258.3   ROW00105           -1.
    COL00116  ROW00118          -.48   ROW00126          -.49
    COL00117  FOB00001         -662.   ROW00106           -1.
    COL00117  ROW00110           -1.   ROW00111           -.5
    COL00117  ROW00112           -.5   ROW00118         -9.32
    COL00117  ROW00125            .7   ROW00126          -.56
    COL00117  ROW00127           -1.   ROW00129            1.
    COL00118  FOB00001            3.   ROW00107           -1.
    COL00118  ROW00115            1.   ROW00126            1.
    COL00119  FOB00001           .39   ROW00108           -1.
    COL00119  ROW00116            1.   ROW00127            1.
    COL00120  FOB00001           .47   ROW00109           -1.
    COL00120  ROW00117            1.   ROW00128            1.
    COL00121  FOB00001          -35.   ROW00111            1.
    COL00122  FOB00001          54.9   ROW00111            1.
    COL00122  ROW00113            1.   ROW00118           -.4
    COL00122  ROW00128           -.5
    COL00123  FOB00001          54.9   ROW00112            1.
    COL00123  ROW00114            1.   ROW00118           -.4
    COL00123  ROW00128           -.5
    COL00124  FOB00001          -35.   ROW00112            1.
    COL00125  FOB00001          23.5   ROW00113           -1.
    COL00125  ROW00118          -1.7   ROW00123            1.
    COL00125  ROW00126         -.245   ROW00128          -1.4
    COL00126  FOB00001          23.5   ROW00114           -1.
    COL00126  ROW00118          -1.7   ROW00124            1.
    COL00126  ROW00126         -.245   ROW00128          -1.4
    COL00127  FOB00001          8.72   ROW00120            1.
    COL00127  ROW00126            .2
    COL00128  FOB00001          9.72   ROW00118           1.5
    COL00128  ROW00119            1.   ROW00120            1.
    COL00129  FOB00001          6.74   ROW00120            1.
    COL00129  ROW00122            1.   ROW00127           1.2
    COL00130  FOB00001          6.84   ROW00120            1.
    COL00130  ROW00122            1.   ROW00128            1.
    COL00131  FOB00001           15.   ROW00118            1.
    COL00132  FOB00001          22.5   ROW00128            1.
    COL00133  ROW00120            1.   ROW00121            1.
    COL00134  FOB00001         -500.   ROW00123           -1.
    COL00135  FOB00001          18.7   ROW00123           -1.
    COL00136  FOB00001        -258.3   ROW00124           -1.
    COL00137  FOB00001         -662.   ROW00125           -1.
    COL00137  ROW00129           -1.
    COL00138  FOB00001            3.   ROW00126           -1.
    COL00139  FOB00001           .39   ROW00127           -1.
    COL00140  FOB00001           .47   ROW00128           -1.
RHS
    RHS       ROW00001          158.   ROW00002          158.
    RHS       ROW00005       3092.96   ROW00006       2566.67
    RHS       ROW00007         6900.   ROW00008         1600.
    RHS       ROW00009          800.   ROW00012        -375.2
    RHS       ROW00013        -92.12   ROW00014         -684.
    RHS       ROW00015         -150.   ROW00020         1800.
    RHS       ROW00021         2400.   ROW00022         1200.
    RHS       ROW00024       2566.67   ROW00025         6900.
    RHS       ROW00026         1600.   ROW00027          800.
    RHS       ROW00039         1800.   ROW00040         2400.
    RHS       ROW00041         1200.   ROW00043       2566.67
    RHS       ROW00044         6900.   ROW00045         1600.
    RHS       ROW00046          800.   ROW00058         1800.
    RHS       ROW00059         2400.   ROW00060         1200.
    RHS       ROW00062       2566.67   ROW00063         6900.
    RHS       ROW00064         1600.   ROW00065          800.
    RHS       ROW00077         1800.   ROW00078         2400.
    RHS       ROW00079         1200.   ROW00081       2566.67
    RHS       ROW00082         6900.   ROW00083         1600.
    RHS       ROW00084          800.   ROW00096         1800.
    RHS       ROW00097         2400.   ROW00098         1200.
    RHS       ROW00100       2566.67   ROW00101         6900.
    RHS       ROW00102         1600.   ROW00103          800.
    RHS       ROW00115         1800.   ROW00116         2400.
    RHS       ROW00117         1200.   ROW00119       2566.67
    RHS       ROW00120         6900.   ROW00121         1600.
    RHS       ROW00122          800.
ENDATA
