NAME          SCAGR25
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
 E  ROW00130
 E  ROW00131
 E  ROW00132
 E  ROW00133
 L  ROW00134
 L  ROW00135
 L  ROW00136
 E  ROW00137
 L  ROW00138
 E  ROW00139
 L  ROW00140
 G  ROW00141
 E  ROW00142
 E  ROW00143
 E  ROW00144
 E  ROW00145
 E  ROW00146
 E  ROW00147
 L  ROW00148
 E  ROW00149
 E  ROW00150
 E  ROW00151
 E  ROW00152
 L  ROW00153
 L  ROW00154
 L  ROW00155
 E  ROW00156
 L  ROW00157
 E  ROW00158
 L  ROW00159
 G  ROW00160
 E  ROW00161
 E  ROW00162
 E  ROW00163
 E  ROW00164
 E  ROW00165
 E  ROW00166
 L  ROW00167
 E  ROW00168
 E  ROW00169
 E  ROW00170
 E  ROW00171
 L  ROW00172
 L  ROW00173
 L  ROW00174
 E  ROW00175
 L  ROW00176
 E  ROW00177
 L  ROW00178
 G  ROW00179
 E  ROW00180
 E  ROW00181
 E  ROW00182
 E  ROW00183
 E  ROW00184
 E  ROW00185
 L  ROW00186
 E  ROW00187
 E  ROW00188
 E  ROW00189
 E  ROW00190
 L  ROW00191
 L  ROW00192
 L  ROW00193
 E  ROW00194
 L  ROW00195
 E  ROW00196
 L  ROW00197
 G  ROW00198
 E  ROW00199
 E  ROW00200
 E  ROW00201
 E  ROW00202
 E  ROW00203
 E  ROW00204
 L  ROW00205
 E  ROW00206
 E  ROW00207
 E  ROW00208
 E  ROW00209
 L  ROW00210
 L  ROW00211
 L  ROW00212
 E  ROW00213
 L  ROW00214
 E  ROW00215
 L  ROW00216
 G  ROW00217
 E  ROW00218
 E  ROW00219
 E  ROW00220
 E  ROW00221
 E  ROW00222
 E  ROW00223
 L  ROW00224
 E  ROW00225
 E  ROW00226
 E  ROW00227
 E  ROW00228
 L  ROW00229
 L  ROW00230
 L  ROW00231
 E  ROW00232
 L  ROW00233
 E  ROW00234
 L  ROW00235
 G  ROW00236
 E  ROW00237
 E  ROW00238
 E  ROW00239
 E  ROW00240
 E  ROW00241
 E  ROW00242
 L  ROW00243
 E  ROW00244
 E  ROW00245
 E  ROW00246
 E  ROW00247
 L  ROW00248
 L  ROW00249
 L  ROW00250
 E  ROW00251
 L  ROW00252
 E  ROW00253
 L  ROW00254
 G  ROW00255
 E  ROW00256
 E  ROW00257
 E  ROW00258
 E  ROW00259
 E  ROW00260
 E  ROW00261
 L  ROW00262
 E  ROW00263
 E  ROW00264
 E  ROW00265
 E  ROW00266
 L  ROW00267
 L  ROW00268
 L  ROW00269
 E  ROW00270
 L  ROW00271
 E  ROW00272
 L  ROW00273
 G  ROW00274
 E  ROW00275
 E  ROW00276
 E  ROW00277
 E  ROW00278
 E  ROW00279
 E  ROW00280
 L  ROW00281
 E  ROW00282
 E  ROW00283
 E  ROW00284
 E  ROW00285
 L  ROW00286
 L  ROW00287
 L  ROW00288
 E  ROW00289
 L  ROW00290
 E  ROW00291
 L  ROW00292
 G  ROW00293
 E  ROW00294
 E  ROW00295
 E  ROW00296
 E  ROW00297
 E  ROW00298
 E  ROW00299
 L  ROW00300
 E  ROW00301
 E  ROW00302
 E  ROW00303
 E  ROW00304
 L  ROW00305
 L  ROW00306
 L  ROW00307
 E  ROW00308
 L  ROW00309
 E  ROW00310
 L  ROW00311
 G  ROW00312
 E  ROW00313
 E  ROW00314
 E  ROW00315
 E  ROW00316
 E  ROW00317
 E  ROW00318
 L  ROW00319
 E  ROW00320
 E  ROW00321
 E  ROW00322
 E  ROW00323
 L  ROW00324
 L  ROW00325
 L  ROW00326
 E  ROW00327
 L  ROW00328
 E  ROW00329
 L  ROW00330
 G  ROW00331
 E  ROW00332
 E  ROW00333
 E  ROW00334
 E  ROW00335
 E  ROW00336
 E  ROW00337
 L  ROW00338
 E  ROW00339
 E  ROW00340
 E  ROW00341
 E  ROW00342
 L  ROW00343
 L  ROW00344
 L  ROW00345
 E  ROW00346
 L  ROW00347
 E  ROW00348
 L  ROW00349
 G  ROW00350
 E  ROW00351
 E  ROW00352
 E  ROW00353
 E  ROW00354
 E  ROW00355
 E  ROW00356
 L  ROW00357
 E  ROW00358
 E  ROW00359
 E  ROW00360
 E  ROW00361
 L  ROW00362
 L  ROW00363
 L  ROW00364
 E  ROW00365
 L  ROW00366
 E  ROW00367
 L  ROW00368
 G  ROW00369
 E  ROW00370
 E  ROW00371
 E  ROW00372
 E  ROW00373
 E  ROW00374
 E  ROW00375
 L  ROW00376
 E  ROW00377
 E  ROW00378
 E  ROW00379
 E  ROW00380
 L  ROW00381
 L  ROW00382
 L  ROW00383
 E  ROW00384
 L  ROW00385
 E  ROW00386
 L  ROW00387
 G  ROW00388
 E  ROW00389
 E  ROW00390
 E  ROW00391
 E  ROW00392
 E  ROW00393
 E  ROW00394
 L  ROW00395
 E  ROW00396
 E  ROW00397
 E  ROW00398
 E  ROW00399
 L  ROW00400
 L  ROW00401
 L  ROW00402
 E  ROW00403
 L  ROW00404
 E  ROW00405
 L  ROW00406
 G  ROW00407
 E  ROW00408
 E  ROW00409
 E  ROW00410
 E  ROW00411
 E  ROW00412
 E  ROW00413
 L  ROW00414
 E  ROW00415
 E  ROW00416
 E  ROW00417
 E  ROW00418
 L  ROW00419
 L  ROW00420
 L  ROW00421
 E  ROW00422
 L  ROW00423
 E  ROW00424
 L  ROW00425
 G  ROW00426
 E  ROW00427
 E  ROW00428
 E  ROW00429
 E  ROW00430
 E  ROW00431
 E  ROW00432
 L  ROW00433
 E  ROW00434
 E  ROW00435
 E  ROW00436
 E  ROW00437
 L  ROW00438
 L  ROW00439
 L  ROW00440
 E  ROW00441
 L  ROW00442
 E  ROW00443
 L  ROW00444
 G  ROW00445
 E  ROW00446
 E  ROW00447
 E  ROW00448
 E  ROW00449
 E  ROW00450
 E  ROW00451
 L  ROW00452
 E  ROW00453
 E  ROW00454
 E  ROW00455
 E  ROW00456
 L  ROW00457
 L  ROW00458
 L  ROW00459
 E  ROW00460
 L  ROW00461
 E  ROW00462
 L  ROW00463
 G  ROW00464
 E  ROW00465
 E  ROW00466
 E  ROW00467
 E  ROW00468
 E  ROW00469
 E  ROW00470
 L  ROW00471
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
    COL00134  ROW00137          -.48   ROW00145          -.49
    COL00135  FOB00001          18.7   ROW00123           -1.
    COL00135  ROW00137          -.48   ROW00144            1.
    COL00135  ROW00145          -.49
    COL00136  FOB00001        -258.3   ROW00124           -1.
    COL00136  ROW00137          -.48   ROW00145          -.49
    COL00137  FOB00001         -662.   ROW00125           -1.
    COL00137  ROW00129           -1.   ROW00130           -.5
    COL00137  ROW00131           -.5   ROW00137         -9.32
    COL00137  ROW00144            .7   ROW00145          -.56
    COL00137  ROW00146           -1.   ROW00148            1.
    COL00138  FOB00001            3.   ROW00126           -1.
    COL00138  ROW00134            1.   ROW00145            1.
    COL00139  FOB00001           .39   ROW00127           -1.
    COL00139  ROW00135            1.   ROW00146            1.
    COL00140  FOB00001           .47   ROW00128           -1.
    COL00140  ROW00136            1.   ROW00147            1.
    COL00141  FOB00001          -35.   ROW00130            1.
    COL00142  FOB00001          54.9   ROW00130            1.
    COL00142  ROW00132            1.   ROW00137           -.4
    COL00142  ROW00147           -.5
    COL00143  FOB00001          54.9   ROW00131            1.
    COL00143  ROW00133            1.   ROW00137           -.4
    COL00143  ROW00147           -.5
    COL00144  FOB00001          -35.   ROW00131            1.
    COL00145  FOB00001          23.5   ROW00132           -1.
    COL00145  ROW00137          -1.7   ROW00142            1.
    COL00145  ROW00145         -.245   ROW00147          -1.4
    COL00146  FOB00001          23.5   ROW00133           -1.
    COL00146  ROW00137          -1.7   ROW00143            1.
    COL00146  ROW00145         -.245   ROW00147          -1.4
    COL00147  FOB00001          8.72   ROW00139            1.
    COL00147  ROW00145            .2
    COL00148  FOB00001          9.72   ROW00137           1.5
    COL00148  ROW00138            1.   ROW00139            1.
    COL00149  FOB00001          6.74   ROW00139            1.
    COL00149  ROW00141            1.   ROW00146           1.2
    COL00150  FOB00001          6.84   ROW00139            1.
    COL00150  ROW00141            1.   ROW00147            1.
    COL00151  FOB00001           15.   ROW00137            1.
    COL00152  FOB00001          22.5   ROW00147            1.
    COL00153  ROW00139            1.   ROW00140            1.
    COL00154  FOB00001         -500.   ROW00142           -1.
    COL00154  ROW00156          -.48   ROW00164          -.49
    COL00155  FOB00001          18.7   ROW00142           -1.
    COL00155  ROW00156          -.48   ROW00163            1.
    COL00155  ROW00164          -.49
    COL00156  FOB00001        -258.3   ROW00143           -1.
    COL00156  ROW00156          -.48   ROW00164          -.49
    COL00157  FOB00001         -662.   ROW00144           -1.
    COL00157  ROW00148           -1.   ROW00149           -.5
    COL00157  ROW00150           -.5   ROW00156         -9.32
    COL00157  ROW00163            .7   ROW00164          -.56
    COL00157  ROW00165           -1.   ROW00167            1.
    COL00158  FOB00001            3.   ROW00145           -1.
    COL00158  ROW00153            1.   ROW00164            1.
    COL00159  FOB00001           .39   ROW00146           -1.
    COL00159  ROW00154            1.   ROW00165            1.
    COL00160  FOB00001           .47   ROW00147           -1.
    COL00160  ROW00155            1.   ROW00166            1.
    COL00161  FOB00001          -35.   ROW00149            1.
    COL00162  FOB00001          54.9   ROW00149            1.
    COL00162  ROW00151            1.   ROW00156           -.4
    COL00162  ROW00166           -.5
    COL00163  FOB00001          54.9   ROW00150            1.
    COL00163  ROW00152            1.   ROW00156           -.4
    COL00163  ROW00166           -.5
    COL00164  FOB00001          -35.   ROW00150            1.
    COL00165  FOB00001          23.5   ROW00151           -1.
    COL00165  ROW00156          -1.7   ROW00161            1.
    COL00165  ROW00164         -.245   ROW00166          -1.4
    COL00166  FOB00001          23.5   ROW00152           -1.
    COL00166  ROW00156          -1.7   ROW00162            1.
    COL00166  ROW00164         -.245   ROW00166          -1.4
    COL00167  FOB00001          8.72   ROW00158            1.
    COL00167  ROW00164            .2
    COL00168  FOB00001          9.72   ROW00156           1.5
    COL00168  ROW00157            1.   ROW00158            1.
    COL00169  FOB00001          6.74   ROW00158            1.
    COL00169  ROW00160            1.   ROW00165           1.2
    COL00170  FOB00001          6.84   ROW00158            1.
    COL00170  ROW00160            1.   ROW00166            1.
    COL00171  FOB00001           15.   ROW00156            1.
    COL00172  FOB00001          22.5   ROW00166            1.
    COL00173  ROW00158            1.   ROW00159            1.
    COL00174  FOB00001         -500.   ROW00161           -1.
    COL00174  ROW00175          -.48   ROW00183          -.49
    COL00175  FOB00001          18.7   ROW00161           -1.
    COL00175  ROW00175          -.48   ROW00182            1.
    COL00175  ROW00183          -.49
    COL00176  FOB00001        -258.3   ROW00162           -1.
    COL00176  ROW00175          -.48   ROW00183          -.49
    COL00177  FOB00001         -662.   ROW00163           -1.
    COL00177  ROW00167           -1.   ROW00168           -.5
    COL00177  ROW00169           -.5   ROW00175         -9.32
    COL00177  ROW00182            .7   ROW00183          -.56
    COL00177  ROW00184           -1.   ROW00186            1.
    COL00178  FOB00001            3.   ROW00164           -1.
    COL00178  ROW00172            1.   ROW00183            1.
    COL00179  FOB00001           .39   ROW00165           -1.
    COL00179  ROW00173            1.   ROW00184            1.
    COL00180  FOB00001           .47   ROW00166           -1.
    COL00180  ROW00174            1.   ROW00185            1.
    COL00181  FOB00001          -35.   ROW00168            1.
    COL00182  FOB00001          54.9   ROW00168            1.
    COL00182  ROW00170            1.   ROW00175           -.4
    COL00182  ROW00185           -.5
    COL00183  FOB00001          54.9   ROW00169            1.
    COL00183  ROW00171            1.   ROW00175           -.4
    COL00183  ROW00185           -.5
    COL00184  FOB00001          -35.   ROW00169            1.
    COL00185  FOB00001          23.5   ROW00170           -1.
    COL00185  ROW00175          -1.7   ROW00180            1.
    COL00185  ROW00183         -.245   ROW00185          -1.4
    COL00186  FOB00001          23.5   ROW00171           -1.
    COL00186  ROW00175          -1.7   ROW00181            1.
    COL00186  ROW00183         -.245   ROW00185          -1.4
    COL00187  FOB00001          8.72   ROW00177            1.
    COL00187  ROW00183            .2
    COL00188  FOB00001          9.72   ROW00175           1.5
    COL00188  ROW00176            1.   ROW00177            1.
    COL00189  FOB00001          6.74   ROW00177            1.
    COL00189  ROW00179            1.   ROW00184           1.2
    COL00190  FOB00001          6.84   ROW00177            1.
    COL00190  ROW00179            1.   ROW00185            1.
    COL00191  FOB00001           15.   ROW00175            1.
    COL00192  FOB00001          22.5   ROW00185            1.
    COL00193  ROW00177            1.   ROW00178            1.
    COL00194  FOB00001         -500.   ROW00180           -1.
    COL00194  ROW00194          -.48   ROW00202          -.49
    COL00195  FOB00001          18.7   ROW00180           -1.
    COL00195  ROW00194          -.48   ROW00201            1.
    COL00195  ROW00202          -.49
    COL00196  FOB00001        -258.3   ROW00181           -1.
    COL00196  ROW00194          -.48   ROW00202          -.49
    COL00197  FOB00001         -662.   ROW00182           -1.
    COL00197  ROW00186           -1.   ROW00187           -.5
    COL00197  ROW00188           -.5   ROW00194         -9.32
    COL00197  ROW00201            .7   ROW00202          -.56
    COL00197  ROW00203           -1.   ROW00205            1.
    COL00198  FOB00001            3.   ROW00183           -1.
    COL00198  ROW00191            1.   ROW00202            1.
    COL00199  FOB00001           .39   ROW00184           -1.
    COL00199  ROW00192            1.   ROW00203            1.
    COL00200  FOB00001           .47   ROW00185           -1.
    COL00200  ROW00193            1.   ROW00204            1.
    COL00201  FOB00001          -35.   ROW00187            1.
    COL00202  FOB00001          54.9   ROW00187            1.
    COL00202  ROW00189            1.   ROW00194           -.4
    COL00202  ROW00204           -.5
    COL00203  FOB00001          54.9   ROW00188            1.
    COL00203  ROW00190            1.   ROW00194           -.4
    COL00203  ROW00204           -.5
    COL00204  FOB00001          -35.   ROW00188            1.
    COL00205  FOB00001          23.5   ROW00189           -1.
    COL00205  ROW00194          -1.7   ROW00199            1.
    COL00205  ROW00202         -.245   ROW00204          -1.4
    COL00206  FOB00001          23.5   ROW00190           -1.
    COL00206  ROW00194          -1.7   ROW00200            1.
    COL00206  ROW00202         -.245   ROW00204          -1.4
    COL00207  FOB00001          8.72   ROW00196            1.
    COL00207  ROW00202            .2
    COL00208  FOB00001          9.72   ROW00194           1.5
    COL00208  ROW00195            1.   ROW00196            1.
    COL00209  FOB00001          6.74   ROW00196            1.
    COL00209  ROW00198            1.   ROW00203           1.2
    COL00210  FOB00001          6.84   ROW00196            1.
    COL00210  ROW00198            1.   ROW00204            1.
    COL00211  FOB00001           15.   ROW00194            1.
    COL00212  FOB00001          22.5   ROW00204            1.
    COL00213  ROW00196            1.   ROW00197            1.
    COL00214  FOB00001         -500.   ROW00199           -1.
    COL00214  ROW00213          -.48   ROW00221          -.49
    COL00215  FOB00001          18.7   ROW00199           -1.
    COL00215  ROW00213          -.48   ROW00220            1.
    COL00215  ROW00221          -.49
    COL00216  FOB00001        -258.3   ROW00200           -1.
    COL00216  ROW00213          -.48   ROW00221          -.49
    COL00217  FOB00001         -662.   ROW00201           -1.
    COL00217  ROW00205           -1.   ROW00206           -.5
    COL00217  ROW00207           -.5   ROW00213         -9.32
    COL00217  ROW00220            .7   ROW00221          -.56
    COL00217  ROW00222           -1.   ROW00224            1.
    COL00218  FOB00001            3.   ROW00202           -1.
    COL00218  ROW00210            1.   ROW00221            1.
    COL00219  FOB00001           .39   ROW00203           -1.
    COL00219  ROW00211            1.   ROW00222            1.
    COL00220  FOB00001           .47   ROW00204           -1.
    COL00220  ROW00212            1.   ROW00223            1.
    COL00221  FOB00001          -35.   ROW00206            1.
    COL00222  FOB00001          54.9   ROW00206            1.
    COL00222  ROW00208            1.   ROW00213           -.4
    COL00222  ROW00223           -.5
    COL00223  FOB00001          54.9   ROW00207            1.
    COL00223  ROW00209            1.   ROW00213           -.4
    COL00223  ROW00223           -.5
    COL00224  FOB00001          -35.   ROW00207            1.
    COL00225  FOB00001          23.5   ROW00208           -1.
    COL00225  ROW00213          -1.7   ROW00218            1.
    COL00225  ROW00221         -.245   ROW00223          -1.4
    COL00226  FOB00001          23.5   ROW00209           -1.
    COL00226  ROW00213          -1.7   ROW00219            1.
    COL00226  ROW00221         -.245   ROW00223          -1.4
    COL00227  FOB00001          8.72   ROW00215            1.
    COL00227  ROW00221            .2
    COL00228  FOB00001          9.72   ROW00213           1.5
    COL00228  ROW00214            1.   ROW00215            1.
    COL00229  FOB00001          6.74   ROW00215            1.
    COL00229  ROW00217            1.   ROW00222           1.2
    COL00230  FOB00001          6.84   ROW00215            1.
    COL00230  ROW00217            1.   ROW00223            1.
    COL00231  FOB00001           15.   ROW00213            1.
    COL00232  FOB00001          22.5   ROW00223            1.
    COL00233  ROW00215            1.   ROW00216            1.
    COL00234  FOB00001         -500.   ROW00218           -1.
    COL00234  ROW00232          -.48   ROW00240          -.49
    COL00235  FOB00001          18.7   ROW00218           -1.
    COL00235  ROW00232          -.48   ROW00239            1.
    COL00235  ROW00240          -.49
    COL00236  FOB00001        -258.3   ROW00219           -1.
    COL00236  ROW00232          -.48   ROW00240          -.49
    COL00237  FOB00001         -662.   ROW00220           -1.
    COL00237  ROW00224           -1.   ROW00225           -.5
    COL00237  ROW00226           -.5   ROW00232         -9.32
    COL00237  ROW00239            .7   ROW00240          -.56
    COL00237  ROW00241           -1.   ROW00243            1.
    COL00238  FOB00001            3.   ROW00221           -1.
    COL00238  ROW00229            1.   ROW00240            1.
    COL00239  FOB00001           .39   ROW00222           -1.
    COL00239  ROW00230            1.   ROW00241            1.
    COL00240  FOB00001           .47   ROW00223           -1.
    COL00240  ROW00231            1.   ROW00242            1.
    COL00241  FOB00001          -35.   ROW00225            1.
    COL00242  FOB00001          54.9   ROW00225            1.
    COL00242  ROW00227            1.   ROW00232           -.4
    COL00242  ROW00242           -.5
    COL00243  FOB00001          54.9   ROW00226            1.
    COL00243  ROW00228            1.   ROW00232           -.4
    COL00243  ROW00242           -.5
    COL00244  FOB00001          -35.   ROW00226            1.
    COL00245  FOB00001          23.5   ROW00227           -1.
    COL00245  ROW00232          -1.7   ROW00237            1.
    COL00245  ROW00240         -.245   ROW00242          -1.4
    COL00246  FOB00001          23.5   ROW00228           -1.
    COL00246  ROW00232          -1.7   ROW00238            1.
    COL00246  ROW00240         -.245   ROW00242          -1.4
    COL00247  FOB00001          8.72   ROW00234            1.
    COL00247  ROW00240            .2
    COL00248  FOB00001          9.72   ROW00232           1.5
    COL00248  ROW00233            1.   ROW00234            1.
    COL00249  FOB00001          6.74   ROW00234            1.
    COL00249  ROW00236            1.   ROW00241           1.2
    COL00250  FOB00001          6.84   ROW00234            1.
    COL00250  ROW00236            1.   ROW00242            1.
    COL00251  FOB00001           15.   ROW00232            1.
    COL00252  FOB00001          22.5   ROW00242            1.
    COL00253  ROW00234            1.   ROW00235            1.
    COL00254  FOB00001         -500.   ROW00237           -1.
    COL00254  ROW00251          -.48   ROW00259          -.49
    COL00255  FOB00001          18.7   ROW00237           -1.
    COL00255  ROW00251          -.48   ROW00258            1.
    COL00255  ROW00259          -.49
    COL00256  FOB00001        -258.3   ROW00238           -1.
    COL00256  ROW00251          -.48   ROW00259          -.49
    COL00257  FOB00001         -662.   ROW00239           -1.
    COL00257  ROW00243           -1.   ROW00244           -.5
    COL00257  ROW00245           -.5   ROW00251         -9.32
    COL00257  ROW00258            .7   ROW00259          -.56
    COL00257  ROW00260           -1.   ROW00262            1.
    COL00258  FOB00001            3.   ROW00240           -1.
    COL00258  ROW00248            1.   ROW00259            1.
    COL00259  FOB00001           .39   ROW00241           -1.
    COL00259  ROW00249            1.   ROW00260            1.
    COL00260  FOB00001           .47   ROW00242           -1.
    COL00260  ROW00250            1.   ROW00261            1.
    COL00261  FOB00001          -35.   ROW00244            1.
    COL00262  FOB00001          54.9   ROW00244            1.
    COL00262  ROW00246            1.   ROW00251           -.4
    COL00262  ROW00261           -.5
    COL00263  FOB00001          54.9   ROW00245            1.
    COL00263  ROW00247            1.   ROW00251           -.4
    COL00263  ROW00261           -.5
    COL00264  FOB00001          -35.   ROW00245            1.
    COL00265  FOB00001          23.5   ROW00246           -1.
    COL00265  ROW00251          -1.7   ROW00256            1.
    COL00265  ROW00259         -.245   ROW00261          -1.4
    COL00266  FOB00001          23.5   ROW00247           -1.
    COL00266  ROW00251          -1.7   ROW00257            1.
    COL00266  ROW00259         -.245   ROW00261          -1.4
    COL00267  FOB00001          8.72   ROW00253            1.
    COL00267  ROW00259            .2
    COL00268  FOB00001          9.72   ROW00251           1.5
    COL00268  ROW00252            1.   ROW00253            1.
    COL00269  FOB00001          6.74   ROW00253            1.
    COL00269  ROW00255            1.   ROW00260           1.2
    COL00270  FOB00001          6.84   ROW00253            1.
    COL00270  ROW00255            1.   ROW00261            1.
    COL00271  FOB00001           15.   ROW00251            1.
    COL00272  FOB00001          22.5   ROW00261            1.
    COL00273  ROW00253            1.   ROW00254            1.
    COL00274  FOB00001         -500.   ROW00256           -1.
    COL00274  ROW00270          -.48   ROW00278          -.49
    COL00275  FOB00001          18.7   ROW00256           -1.
    COL00275  ROW00270          -.48   ROW00277            1.
    COL00275  ROW00278          -.49
    COL00276  FOB00001        -258.3   ROW00257           -1.
    COL00276  ROW00270          -.48   ROW00278          -.49
    COL00277  FOB00001         -662.   ROW00258           -1.
    COL00277  ROW00262           -1.   ROW00263           -.5
    COL00277  ROW00264           -.5   ROW00270         -9.32
    COL00277  ROW00277            .7   ROW00278          -.56
    COL00277  ROW00279           -1.   ROW00281            1.
    COL00278  FOB00001            3.   ROW00259           -1.
    COL00278  ROW00267            1.   ROW00278            1.
    COL00279  FOB00001           .39   ROW00260           -1.
    COL00279  ROW00268            1.   ROW00279            1.
    COL00280  FOB00001           .47   ROW00261           -1.
    COL00280  ROW00269            1.   ROW00280            1.
    COL00281  FOB00001          -35.   ROW00263            1.
    COL00282  FOB00001          54.9   ROW00263            1.
    COL00282  ROW00265            1.   ROW00270           -.4
    COL00282  ROW00280           -.5
    COL00283  FOB00001          54.9   ROW00264            1.
    COL00283  ROW00266            1.   ROW00270           -.4
    COL00283  ROW00280           -.5
    COL00284  FOB00001          -35.   ROW00264            1.
    COL00285  FOB00001          23.5   ROW00265           -1.
    COL00285  ROW00270          -1.7   ROW00275            1.
    COL00285  ROW00278         -.245   ROW00280          -1.4
    COL00286  FOB00001          23.5   ROW00266           -1.
    COL00286  ROW00270          -1.7   ROW00276            1.
    COL00286  ROW00278         -.245   ROW00280          -1.4
    COL00287  FOB00001          8.72   ROW00272            1.
    COL00287  ROW00278            .2
    COL00288  FOB00001          9.72   ROW00270           1.5
    COL00288  ROW00271            1.   ROW00272            1.
    COL00289  FOB00001          6.74   ROW00272            1.
    COL00289  ROW00274            1.   ROW00279           1.2
    COL00290  FOB00001          6.84   ROW00272            1.
    COL00290  ROW00274            1.   ROW00280            1.
    COL00291  FOB00001           15.   ROW00270            1.
    COL00292  FOB00001          22.5   ROW00280            1.
    COL00293  ROW00272            1.   ROW00273            1.
    COL00294  FOB00001         -500.   ROW00275           -1.
    COL00294  ROW00289          -.48   ROW00297          -.49
    COL00295  FOB00001          18.7   ROW00275           -1.
    COL00295  ROW00289          -.48   ROW00296            1.
    COL00295  ROW00297          -.49
    COL00296  FOB00001        -258.3   ROW00276           -1.
    COL00296  ROW00289          -.48   ROW00297          -.49
    COL00297  FOB00001         -662.   ROW00277           -1.
    COL00297  ROW00281           -1.   ROW00282           -.5
    COL00297  ROW00283           -.5   ROW00289         -9.32
    COL00297  ROW00296            .7   ROW00297          -.56
    COL00297  ROW00298           -1.   ROW00300            1.
    COL00298  FOB00001            3.   ROW00278           -1.
    COL00298  ROW00286            1.   ROW00297            1.
    COL00299  FOB00001           .39   ROW00279           -1.
    COL00299  ROW00287            1.   ROW00298            1.
    COL00300  FOB00001           .47   ROW00280           -1.
    COL00300  ROW00288            1.   ROW00299            1.
    COL00301  FOB00001          -35.   ROW00282            1.
    COL00302  FOB00001          54.9   ROW00282            1.
    COL00302  ROW00284            1.   ROW00289           -.4
    COL00302  ROW00299           -.5
    COL00303  FOB00001          54.9   ROW00283            1.
    COL00303  ROW00285            1.   ROW00289           -.4
    COL00303  ROW00299           -.5
    COL00304  FOB00001          -35.   ROW00283            1.
    COL00305  FOB00001          23.5   ROW00284           -1.
    COL00305  ROW00289          -1.7   ROW00294            1.
    COL00305  ROW00297         -.245   ROW00299          -1.4
    COL00306  FOB00001          23.5   ROW00285           -1.
    COL00306  ROW00289          -1.7   ROW00295            1.
    COL00306  ROW00297         -.245   ROW00299          -1.4
    COL00307  FOB00001          8.72   ROW00291            1.
    COL00307  ROW00297            .2
    COL00308  FOB00001          9.72   ROW00289           1.5
    COL00308  ROW00290            1.   ROW00291            1.
    COL00309  FOB00001          6.74   ROW00291            1.
    COL00309  ROW00293            1.   ROW00298           1.2
    COL00310  FOB00001          6.84   ROW00291            1.
    COL00310  ROW00293            1.   ROW00299            1.
    COL00311  FOB00001           15.   ROW00289            1.
    COL00312  FOB00001          22.5   ROW00299            1.
    COL00313  ROW00291            1.   ROW00292            1.
    COL00314  FOB00001         -500.   ROW00294           -1.
    COL00314  ROW00308          -.48   ROW00316          -.49
    COL00315  FOB00001          18.7   ROW00294           -1.
    COL00315  ROW00308          -.48   ROW00315            1.
    COL00315  ROW00316          -.49
    COL00316  FOB00001        -258.3   ROW00295           -1.
    COL00316  ROW00308          -.48   ROW00316          -.49
    COL00317  FOB00001         -662.   ROW00296           -1.
    COL00317  ROW00300           -1.   ROW00301           -.5
    COL00317  ROW00302           -.5   ROW00308         -9.32
    COL00317  ROW00315            .7   ROW00316          -.56
    COL00317  ROW00317           -1.   ROW00319            1.
    COL00318  FOB00001            3.   ROW00297           -1.
    COL00318  ROW00305            1.   ROW00316            1.
    COL00319  FOB00001           .39   ROW00298           -1.
    COL00319  ROW00306            1.   ROW00317            1.
    COL00320  FOB00001           .47   ROW00299           -1.
    COL00320  ROW00307            1.   ROW00318            1.
    COL00321  FOB00001          -35.   ROW00301            1.
    COL00322  FOB00001          54.9   ROW00301            1.
    COL00322  ROW00303            1.   ROW00308           -.4
    COL00322  ROW00318           -.5
    COL00323  FOB00001          54.9   ROW00302            1.
    COL00323  ROW00304            1.   ROW00308           -.4
    COL00323  ROW00318           -.5
    COL00324  FOB00001          -35.   ROW00302            1.
    COL00325  FOB00001          23.5   ROW00303           -1.
    COL00325  ROW00308          -1.7   ROW00313            1.
    COL00325  ROW00316         -.245   ROW00318          -1.4
    COL00326  FOB00001          23.5   ROW00304           -1.
    COL00326  ROW00308          -1.7   ROW00314            1.
    COL00326  ROW00316         -.245   ROW00318          -1.4
    COL00327  FOB00001          8.72   ROW00310            1.
    COL00327  ROW00316            .2
    COL00328  FOB00001          9.72   ROW00308           1.5
    COL00328  ROW00309            1.   ROW00310            1.
    COL00329  FOB00001          6.74   ROW00310            1.
    COL00329  ROW00312            1.   ROW00317           1.2
    COL00330  FOB00001          6.84   ROW00310            1.
    COL00330  ROW00312            1.   ROW00318            1.
    COL00331  FOB00001           15.   ROW00308            1.
    COL00332  FOB00001          22.5   ROW00318            1.
    COL00333  ROW00310            1.   ROW00311            1.
    COL00334  FOB00001         -500.   ROW00313           -1.
    COL00334  ROW00327          -.48   ROW00335          -.49
    COL00335  FOB00001          18.7   ROW00313           -1.
    COL00335  ROW00327          -.48   ROW00334            1.
    COL00335  ROW00335          -.49
    COL00336  FOB00001        -258.3   ROW00314           -1.
    COL00336  ROW00327          -.48   ROW00335          -.49
    COL00337  FOB00001         -662.   ROW00315           -1.
    COL00337  ROW00319           -1.   ROW00320           -.5
    COL00337  ROW00321           -.5   ROW00327         -9.32
    COL00337  ROW00334            .7   ROW00335          -.56
    COL00337  ROW00336           -1.   ROW00338            1.
    COL00338  FOB00001            3.   ROW00316           -1.
    COL00338  ROW00324            1.   ROW00335            1.
    COL00339  FOB00001           .39   ROW00317           -1.
    COL00339  ROW00325            1.   ROW00336            1.
    COL00340  FOB00001           .47   ROW00318           -1.
    COL00340  ROW00326            1.   ROW00337            1.
    COL00341  FOB00001          -35.   ROW00320            1.
    COL00342  FOB00001          54.9   ROW00320            1.
    COL00342  ROW00322            1.   ROW00327           -.4
    COL00342  ROW00337           -.5
    COL00343  FOB00001          54.9   ROW00321            1.
    COL00343  ROW00323            1.   ROW00327           -.4
    COL00343  ROW00337           -.5
    COL00344  FOB00001          -35.   ROW00321            1.
    COL00345  FOB00001          23.5   ROW00322           -1.
    COL00345  ROW00327          -1.7   ROW00332            1.
    COL00345  ROW00335         -.245   ROW00337          -1.4
    COL00346  FOB00001          23.5   ROW00323           -1.
    COL00346  ROW00327          -1.7   ROW00333            1.
    COL00346  ROW00335         -.245   ROW00337          -1.4
    COL00347  FOB00001          8.72   ROW00329            1.
    COL00347  ROW00335            .2
    COL00348  FOB00001          9.72   ROW00327           1.5
    COL00348  ROW00328            1.   ROW00329            1.
    COL00349  FOB00001          6.74   ROW00329            1.
    COL00349  ROW00331            1.   ROW00336           1.2
    COL00350  FOB00001          6.84   ROW00329            1.
    COL00350  ROW00331            1.   ROW00337            1.
    COL00351  FOB00001           15.   ROW00327            1.
    COL00352  FOB00001          22.5   ROW00337            1.
    COL00353  ROW00329            1.   ROW00330            1.
    COL00354  FOB00001         -500.   ROW00332           -1.
    COL00354  ROW00346          -.48   ROW00354          -.49
    COL00355  FOB00001          18.7   ROW00332           -1.
    COL00355  ROW00346          -.48   ROW00353            1.
    COL00355  ROW00354          -.49
    COL00356  FOB00001        -258.3   ROW00333           -1.
    COL00356  ROW00346          -.48   ROW00354          -.49
    COL00357  FOB00001         -662.   ROW00334           -1.
    COL00357  ROW00338           -1.   ROW00339           -.5
    COL00357  ROW00340           -.5   ROW00346         -9.32
    COL00357  ROW00353            .7   ROW00354          -.56
    COL00357  ROW00355           -1.   ROW00357            1.
    COL00358  FOB00001            3.   ROW00335           -1.
    COL00358  ROW00343            1.   ROW00354            1.
    COL00359  FOB00001           .39   ROW00336           -1.
    COL00359  ROW00344            1.   ROW00355            1.
    COL00360  FOB00001           .47   ROW00337           -1.
    COL00360  ROW00345            1.   ROW00356            1.
    COL00361  FOB00001          -35.   ROW00339            1.
    COL00362  FOB00001          54.9   ROW00339            1.
    COL00362  ROW00341            1.   ROW00346           -.4
    COL00362  ROW00356           -.5
    COL00363  FOB00001          54.9   ROW00340            1.
    COL00363  ROW00342            1.   ROW00346           -.4
    COL00363  ROW00356           -.5
    COL00364  FOB00001          -35.   ROW00340            1.
    COL00365  FOB00001          23.5   ROW00341           -1.
    COL00365  ROW00346          -1.7   ROW00351            1.
    COL00365  ROW00354         -.245   ROW00356          -1.4
    COL00366  FOB00001          23.5   ROW00342           -1.
    COL00366  ROW00346          -1.7   ROW00352            1.
    COL00366  ROW00354         -.245   ROW00356          -1.4
    COL00367  FOB00001          8.72   ROW00348            1.
    COL00367  ROW00354            .2
    COL00368  FOB00001          9.72   ROW00346           1.5
    COL00368  ROW00347            1.   ROW00348            1.
    COL00369  FOB00001          6.74   ROW00348            1.
    COL00369  ROW00350            1.   ROW00355           1.2
    COL00370  FOB00001          6.84   ROW00348            1.
    COL00370  ROW00350            1.   ROW00356            1.
    COL00371  FOB00001           15.   ROW00346            1.
    COL00372  FOB00001          22.5   ROW00356            1.
    COL00373  ROW00348            1.   ROW00349            1.
    COL00374  FOB00001         -500.   ROW00351           -1.
    COL00374  ROW00365          -.48   ROW00373          -.49
    COL00375  FOB00001          18.7   ROW00351           -1.
    COL00375  ROW00365          -.48   ROW00372            1.
    COL00375  ROW00373          -.49
    COL00376  FOB00001        -258.3   ROW00352           -1.
    COL00376  ROW00365          -.48   ROW00373          -.49
    COL00377  FOB00001         -662.   ROW00353           -1.
    COL00377  ROW00357           -1.   ROW00358           -.5
    COL00377  ROW00359           -.5   ROW00365         -9.32
    COL00377  ROW00372            .7   ROW00373          -.56
    COL00377  ROW00374           -1.   ROW00376            1.
    COL00378  FOB00001            3.   ROW00354           -1.
    COL00378  ROW00362            1.   ROW00373            1.
    COL00379  FOB00001           .39   ROW00355           -1.
    COL00379  ROW00363            1.   ROW00374            1.
    COL00380  FOB00001           .47   ROW00356           -1.
    COL00380  ROW00364            1.   ROW00375            1.
    COL00381  FOB00001          -35.   ROW00358            1.
    COL00382  FOB00001          54.9   ROW00358            1.
    COL00382  ROW00360            1.   ROW00365           -.4
    COL00382  ROW00375           -.5
    COL00383  FOB00001          54.9   ROW00359            1.
    COL00383  ROW00361            1.   ROW00365           -.4
    COL00383  ROW00375           -.5
    COL00384  FOB00001          -35.   ROW00359            1.
    COL00385  FOB00001          23.5   ROW00360           -1.
    COL00385  ROW00365          -1.7   ROW00370            1.
    COL00385  ROW00373         -.245   ROW00375          -1.4
    COL00386  FOB00001          23.5   ROW00361           -1.
    COL00386  ROW00365          -1.7   ROW00371            1.
    COL00386  ROW00373         -.245   ROW00375          -1.4
    COL00387  FOB00001          8.72   ROW00367            1.
    COL00387  ROW00373            .2
    COL00388  FOB00001          9.72   ROW00365           1.5
    COL00388  ROW00366            1.   ROW00367            1.
    COL00389  FOB00001          6.74   ROW00367            1.
    COL00389  ROW00369            1.   ROW00374           1.2
    COL00390  FOB00001          6.84   ROW00367            1.
    COL00390  ROW00369            1.   ROW00375            1.
    COL00391  FOB00001           15.   ROW00365            1.
    COL00392  FOB00001          22.5   ROW00375            1.
    COL00393  ROW00367            1.   ROW00368            1.
    COL00394  FOB00001         -500.   ROW00370           -1.
    COL00394  ROW00384          -.48   ROW00392          -.49
    COL00395  FOB00001          18.7   ROW00370           -1.
    COL00395  ROW00384          -.48   ROW00391            1.
    COL00395  ROW00392          -.49
    COL00396  FOB00001        -258.3   ROW00371           -1.
    COL00396  ROW00384          -.48   ROW00392          -.49
    COL00397  FOB00001         -662.   ROW00372           -1.
    COL00397  ROW00376           -1.   ROW00377           -.5
    COL00397  ROW00378           -.5   ROW00384         -9.32
    COL00397  ROW00391            .7   ROW00392          -.56
    COL00397  ROW00393           -1.   ROW00395            1.
    COL00398  FOB00001            3.   ROW00373           -1.
    COL00398  ROW00381            1.   ROW00392            1.
    COL00399  FOB00001           .39   ROW00374           -1.
    COL00399  ROW00382            1.   ROW00393            1.
    COL00400  FOB00001           .47   ROW00375           -1.
    COL00400  ROW00383            1.   ROW00394            1.
    COL00401  FOB00001          -35.   ROW00377            1.
    COL00402  FOB00001          54.9   ROW00377            1.
    COL00402  ROW00379            1.   ROW00384           -.4
    COL00402  ROW00394           -.5
    COL00403  FOB00001          54.9   ROW00378            1.
    COL00403  ROW00380            1.   ROW00384           -.4
    COL00403  ROW00394           -.5
    COL00404  FOB00001          -35.   ROW00378            1.
    COL00405  FOB00001          23.5   ROW00379           -1.
    COL00405  ROW00384          -1.7   ROW00389            1.
    COL00405  ROW00392         -.245   ROW00394          -1.4
    COL00406  FOB00001          23.5   ROW00380           -1.
    COL00406  ROW00384          -1.7   ROW00390            1.
    COL00406  ROW00392         -.245   ROW00394          -1.4
    COL00407  FOB00001          8.72   ROW00386            1.
    COL00407  ROW00392            .2
    COL00408  FOB00001          9.72   ROW00384           1.5
    COL00408  ROW00385            1.   ROW00386            1.
    COL00409  FOB00001          6.74   ROW00386            1.
    COL00409  ROW00388            1.   ROW00393           1.2
    COL00410  FOB00001          6.84   ROW00386            1.
    COL00410  ROW00388            1.   ROW00394            1.
    COL00411  FOB00001           15.   ROW00384            1.
    COL00412  FOB00001          22.5   ROW00394            1.
    COL00413  ROW00386            1.   ROW00387            1.
    COL00414  FOB00001         -500.   ROW00389           -1.
    COL00414  ROW00403          -.48   ROW00411          -.49
    COL00415  FOB00001          18.7   ROW00389           -1.
    COL00415  ROW00403          -.48   ROW00410            1.
    COL00415  ROW00411          -.49
    COL00416  FOB00001        -258.3   ROW00390           -1.
    COL00416  ROW00403          -.48   ROW00411          -.49
    COL00417  FOB00001         -662.   ROW00391           -1.
    COL00417  ROW00395           -1.   ROW00396           -.5
    COL00417  ROW00397           -.5   ROW00403         -9.32
    COL00417  ROW00410            .7   ROW00411          -.56
    COL00417  ROW00412           -1.   ROW00414            1.
    COL00418  FOB00001            3.   ROW00392           -1.
    COL00418  ROW00400            1.   ROW00411            1.
    COL00419  FOB00001           .39   ROW00393           -1.
    COL00419  ROW00401            1.   ROW00412            1.
    COL00420  FOB00001           .47   ROW00394           -1.
    COL00420  ROW00402            1.   ROW00413            1.
    COL00421  FOB00001          -35.   ROW00396            1.
    COL00422  FOB00001          54.9   ROW00396            1.
    COL00422  ROW00398            1.   ROW00403           -.4
    COL00422  ROW00413           -.5
    COL00423  FOB00001          54.9   ROW00397            1.
    COL00423  ROW00399            1.   ROW00403           -.4
    COL00423  ROW00413           -.5
    COL00424  FOB00001          -35.   ROW00397            1.
    COL00425  FOB00001          23.5   ROW00398           -1.
    COL00425  ROW00403          -1.7   ROW00408            1.
    COL00425  ROW00411         -.245   ROW00413          -1.4
    COL00426  FOB00001          23.5   ROW00399           -1.
    COL00426  ROW00403          -1.7   ROW00409            1.
    COL00426  ROW00411         -.245   ROW00413          -1.4
    COL00427  FOB00001          8.72   ROW00405            1.
    COL00427  ROW00411            .2
    COL00428  FOB00001          9.72   ROW00403           1.5
    COL00428  ROW00404            1.   ROW00405            1.
    COL00429  FOB00001          6.74   ROW00405            1.
    COL00429  ROW00407            1.   ROW00412           1.2
    COL00430  FOB00001          6.84   ROW00405            1.
    COL00430  ROW00407            1.   ROW00413            1.
    COL00431  FOB00001           15.   ROW00403            1.
    COL00432  FOB00001          22.5   ROW00413            1.
    COL00433  ROW00405            1.   ROW00406            1.
    COL00434  FOB00001         -500.   ROW00408           -1.
    COL00434  ROW00422          -.48   ROW00430          -.49
    COL00435  FOB00001          18.7   ROW00408           -1.
    COL00435  ROW00422          -.48   ROW00429            1.
    COL00435  ROW00430          -.49
    COL00436  FOB00001        -258.3   ROW00409           -1.
    COL00436  ROW00422          -.48   ROW00430          -.49
    COL00437  FOB00001         -662.   ROW00410           -1.
    COL00437  ROW00414           -1.   ROW00415           -.5
    COL00437  ROW00416           -.5   ROW00422         -9.32
    COL00437  ROW00429            .7   ROW00430          -.56
    COL00437  ROW00431           -1.   ROW00433            1.
    COL00438  FOB00001            3.   ROW00411           -1.
    COL00438  ROW00419            1.   ROW00430            1.
    COL00439  FOB00001           .39   ROW00412           -1.
    COL00439  ROW00420            1.   ROW00431            1.
    COL00440  FOB00001           .47   ROW00413           -1.
    COL00440  ROW00421            1.   ROW00432            1.
    COL00441  FOB00001          -35.   ROW00415            1.
    COL00442  FOB00001          54.9   ROW00415            1.
    COL00442  ROW00417            1.   ROW00422           -.4
    COL00442  ROW00432           -.5
    COL00443  FOB00001          54.9   ROW00416            1.
    COL00443  ROW00418            1.   ROW00422           -.4
    COL00443  ROW00432           -.5
    COL00444  FOB00001          -35.   ROW00416            1.
    COL00445  FOB00001          23.5   ROW00417           -1.
    COL00445  ROW00422          -1.7   ROW00427            1.
    COL00445  ROW00430         -.245   ROW00432          -1.4
    COL00446  FOB00001          23.5   ROW00418           -1.
    COL00446  ROW00422          -1.7   ROW00428            1.
    COL00446  ROW00430         -.245   ROW00432          -1.4
    COL00447  FOB00001          8.72   ROW00424            1.
    COL00447  ROW00430            .2
    COL00448  FOB00001          9.72   ROW00422           1.5
    COL00448  ROW00423            1.   ROW00424            1.
    COL00449  FOB00001          6.74   ROW00424            1.
    COL00449  ROW00426            1.   ROW00431           1.2
    COL00450  FOB00001          6.84   ROW00424            1.
    COL00450  ROW00426            1.   ROW00432            1.
    COL00451  FOB00001           15.   ROW00422            1.
    COL00452  FOB00001          22.5   ROW00432            1.
    COL00453  ROW00424            1.   ROW00425            1.
    COL00454  FOB00001         -500.   ROW00427           -1.
    COL00454  ROW00441          -.48   ROW00449          -.49
    COL00455  FOB00001          18.7   ROW00427           -1.
    COL00455  ROW00441          -.48   ROW00448            1.
    COL00455  ROW00449          -.49
    COL00456  FOB00001        -258.3   ROW00428           -1.
    COL00456  ROW00441          -.48   ROW00449          -.49
    COL00457  FOB00001         -662.   ROW00429           -1.
    COL00457  ROW00433           -1.   ROW00434           -.5
    COL00457  ROW00435           -.5   ROW00441         -9.32
    COL00457  ROW00448            .7   ROW00449          -.56
    COL00457  ROW00450           -1.   ROW00452            1.
    COL00458  FOB00001            3.   ROW00430           -1.
    COL00458  ROW00438            1.   ROW00449            1.
    COL00459  FOB00001           .39   ROW00431           -1.
    COL00459  ROW00439            1.   ROW00450            1.
    COL00460  FOB00001           .47   ROW00432           -1.
    COL00460  ROW00440            1.   ROW00451            1.
    COL00461  FOB00001          -35.   ROW00434            1.
    COL00462  FOB00001          54.9   ROW00434            1.
    COL00462  ROW00436            1.   ROW00441           -.4
    COL00462  ROW00451           -.5
    COL00463  FOB00001          54.9   ROW00435            1.
    COL00463  ROW00437            1.   ROW00441           -.4
    COL00463  ROW00451           -.5
    COL00464  FOB00001          -35.   ROW00435            1.
    COL00465  FOB00001          23.5   ROW00436           -1.
    COL00465  ROW00441          -1.7   ROW00446            1.
    COL00465  ROW00449         -.245   ROW00451          -1.4
    COL00466  FOB00001          23.5   ROW00437           -1.
    COL00466  ROW00441          -1.7   ROW00447            1.
    COL00466  ROW00449         -.245   ROW00451          -1.4
    COL00467  FOB00001          8.72   ROW00443            1.
    COL00467  ROW00449            .2
    COL00468  FOB00001          9.72   ROW00441           1.5
    COL00468  ROW00442            1.   ROW00443            1.
    COL00469  FOB00001          6.74   ROW00443            1.
    COL00469  ROW00445            1.   ROW00450           1.2
    COL00470  FOB00001          6.84   ROW00443            1.
    COL00470  ROW00445            1.   ROW00451            1.
    COL00471  FOB00001           15.   ROW00441            1.
    COL00472  FOB00001          22.5   ROW00451            1.
    COL00473  ROW00443            1.   ROW00444            1.
    COL00474  FOB00001         -500.   ROW00446           -1.
    COL00474  ROW00460          -.48   ROW00468          -.49
    COL00475  FOB00001          18.7   ROW00446           -1.
    COL00475  ROW00460          -.48   ROW00467            1.
    COL00475  ROW00468          -.49
    COL00476  FOB00001        -258.3   ROW00447           -1.
    COL00476  ROW00460          -.48   ROW00468          -.49
    COL00477  FOB00001         -662.   ROW00448           -1.
    COL00477  ROW00452           -1.   ROW00453           -.5
    COL00477  ROW00454           -.5   ROW00460         -9.32
    COL00477  ROW00467            .7   ROW00468          -.56
    COL00477  ROW00469           -1.   ROW00471            1.
    COL00478  FOB00001            3.   ROW00449           -1.
    COL00478  ROW00457            1.   ROW00468            1.
    COL00479  FOB00001           .39   ROW00450           -1.
    COL00479  ROW00458            1.   ROW00469            1.
    COL00480  FOB00001           .47   ROW00451           -1.
    COL00480  ROW00459            1.   ROW00470            1.
    COL00481  FOB00001          -35.   ROW00453            1.
    COL00482  FOB00001          54.9   ROW00453            1.
    COL00482  ROW00455            1.   ROW00460           -.4
    COL00482  ROW00470           -.5
    COL00483  FOB00001          54.9   ROW00454            1.
    COL00483  ROW00456            1.   ROW00460           -.4
    COL00483  ROW00470           -.5
    COL00484  FOB00001          -35.   ROW00454            1.
    COL00485  FOB00001          23.5   ROW00455           -1.
    COL00485  ROW00460          -1.7   ROW00465            1.
    COL00485  ROW00468         -.245   ROW00470          -1.4
    COL00486  FOB00001          23.5   ROW00456           -1.
    COL00486  ROW00460          -1.7   ROW00466            1.
    COL00486  ROW00468         -.245   ROW00470          -1.4
    COL00487  FOB00001          8.72   ROW00462            1.
    COL00487  ROW00468            .2
    COL00488  FOB00001          9.72   ROW00460           1.5
    COL00488  ROW00461            1.   ROW00462            1.
    COL00489  FOB00001          6.74   ROW00462            1.
    COL00489  ROW00464            1.   ROW00469           1.2
    COL00490  FOB00001          6.84   ROW00462            1.
    COL00490  ROW00464            1.   ROW00470            1.
    COL00491  FOB00001           15.   ROW00460            1.
    COL00492  FOB00001          22.5   ROW00470            1.
    COL00493  ROW00462            1.   ROW00463            1.
    COL00494  FOB00001         -500.   ROW00465           -1.
    COL00495  FOB00001          18.7   ROW00465           -1.
    COL00496  FOB00001        -258.3   ROW00466           -1.
    COL00497  FOB00001         -662.   ROW00467           -1.
    COL00497  ROW00471           -1.
    COL00498  FOB00001            3.   ROW00468           -1.
    COL00499  FOB00001           .39   ROW00469           -1.
    COL00500  FOB00001           .47   ROW00470           -1.
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
    RHS       ROW00122          800.   ROW00134         1800.
    RHS       ROW00135         2400.   ROW00136         1200.
    RHS       ROW00138       2566.67   ROW00139         6900.
    RHS       ROW00140         1600.   ROW00141          800.
    RHS       ROW00153         1800.   ROW00154         2400.
    RHS       ROW00155         1200.   ROW00157       2566.67
    RHS       ROW00158         6900.   ROW00159         1600.
    RHS       ROW00160          800.   ROW00172         1800.
    RHS       ROW00173         2400.   ROW00174         1200.
    RHS       ROW00176       2566.67   ROW00177         6900.
    RHS       ROW00178         1600.   ROW00179          800.
    RHS       ROW00191         1800.   ROW00192         2400.
    RHS       ROW00193         1200.   ROW00195       2566.67
    RHS       ROW00196         6900.   ROW00197         1600.
    RHS       ROW00198          800.   ROW00210         1800.
    RHS       ROW00211         2400.   ROW00212         1200.
    RHS       ROW00214       2566.67   ROW00215         6900.
    RHS       ROW00216         1600.   ROW00217          800.
    RHS       ROW00229         1800.   ROW00230         2400.
    RHS       ROW00231         1200.   ROW00233       2566.67
    RHS       ROW00234         6900.   ROW00235         1600.
    RHS       ROW00236          800.   ROW00248         1800.
    RHS       ROW00249         2400.   ROW00250         1200.
    RHS       ROW00252       2566.67   ROW00253         6900.
    RHS       ROW00254         1600.   ROW00255          800.
    RHS       ROW00267         1800.   ROW00268         2400.
    RHS       ROW00269         1200.   ROW00271       2566.67
    RHS       ROW00272         6900.   ROW00273         1600.
    RHS       ROW00274          800.   ROW00286         1800.
    RHS       ROW00287         2400.   ROW00288         1200.
    RHS       ROW00290       2566.67   ROW00291         6900.
    RHS       ROW00292         1600.   ROW00293          800.
    RHS       ROW00305         1800.   ROW00306         2400.
    RHS       ROW00307         1200.   ROW00309       2566.67
    RHS       ROW00310         6900.   ROW00311         1600.
    RHS       ROW00312          800.   ROW00324         1800.
    RHS       ROW00325         2400.   ROW00326         1200.
    RHS       ROW00328       2566.67   ROW00329         6900.
    RHS       ROW00330         1600.   ROW00331          800.
    RHS       ROW00343         1800.   ROW00344         2400.
    RHS       ROW00345         1200.   ROW00347       2566.67
    RHS       ROW00348         6900.   ROW00349         1600.
    RHS       ROW00350          800.   ROW00362         1800.
    RHS       ROW00363         2400.   ROW00364         1200.
    RHS       ROW00366       2566.67   ROW00367         6900.
    RHS       ROW00368         1600.   ROW00369          800.
    RHS       ROW00381         1800.   ROW00382         2400.
    RHS       ROW00383         1200.   ROW00385       2566.67
    RHS       ROW00386         6900.   ROW00387         1600.
    RHS       ROW00388          800.   ROW00400         1800.
    RHS       ROW00401         2400.   ROW00402         1200.
    RHS       ROW00404       2566.67   ROW00405         6900.
    RHS       ROW00406         1600.   ROW00407          800.
    RHS       ROW00419         1800.   ROW00420         2400.
    RHS       ROW00421         1200.   ROW00423       2566.67
    RHS       ROW00424         6900.   ROW00425         1600.
    RHS       ROW00426          800.   ROW00438         1800.
    RHS       ROW00439         2400.   ROW00440         1200.
    RHS       ROW00442       2566.67   ROW00443         6900.
    RHS       ROW00444         1600.   ROW00445          800.
    RHS       ROW00457         1800.   ROW00458         2400.
    RHS       ROW00459         1200.   ROW00461       2566.67
    RHS       ROW00462         6900.   ROW00463         1600.
    RHS       ROW00464          800.
ENDATA
