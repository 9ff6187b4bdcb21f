NAME          CAPRI
ROWS
 E  VLRES
 N  OBJEC
 E  RAI72
 G  CVI72
 G  TRS72
 E  FPR72
 E  DEP72
 E  DEP73
 E  INV72
 E  INC72
 E  INT72
 G  TRS73
 G  CAF72
 E  FPR73
 E  RAI73
 G  CVI73
 E  DEP74
 E  INV73
 E  INC73
 E  INT73
 G  TRS74
 G  CAF73
 E  FPR74
 E  RAI74
 G  CVI74
 E  DEP75
 E  INV74
 E  INC74
 E  INT74
 G  TRS75
 G  CAF74
 E  FPR75
 E  RAI75
 G  CVI75
 E  DEP76
 E  INV75
 E  INC75
 E  INT75
 G  TRS76
 G  CAF75
 E  FPR76
 E  RAI76
 G  CVI76
 E  DEP77
 E  INV76
 E  INC76
 E  INT76
 G  TRS77
 G  CAF76
 E  FPR77
 E  RAI77
 G  CVI77
 E  DEP78
 E  INV77
 E  INC77
 E  INT77
 G  TRS78
 G  CAF77
 E  FPR78
 E  RAI78
 G  CVI78
 E  DEP79
 E  INV78
 E  INC78
 E  INT78
 G  TRS79
 G  CAF78
 E  FPR79
 E  RAI79
 G  CVI79
 E  DEP80
 E  INV79
 E  INC79
 E  INT79
 G  TRS80
 G  CAF79
 E  FPR80
 E  RAI80
 G  CVI80
 E  DEP81
 E  INV80
 E  INC80
 E  INT80
 G  TRS81
 G  CAF80
 E  FPR81
 E  RAI81
 G  CVI81
 E  INV81
 E  INC81
 E  INT81
 G  CAF81
 E  L1378
 E  L1379
 L  R1378
 E  FXE78
 E  FXE79
 E  FXE80
 E  FXE81
 E  XX178
 E  XX179
 E  XX378
 E  XX379
 E  L2378
 E  L2379
 L  R2378
 E  XX278
 E  XX279
 E  L3578
 E  L3579
 L  R3578
 E  XX578
 E  XX579
 E  L5478
 E  L5479
 L  R5478
 E  L5678
 E  L5679
 L  R5678
 E  L1380
 L  R1379
 E  XX180
 E  XX380
 E  L2380
 L  R2379
 E  XX280
 E  L3580
 L  R3579
 E  XX580
 E  L5480
 L  R5479
 E  L5680
 L  R5679
 E  L1381
 L  R1380
 E  XX181
 E  XX381
 E  L2381
 L  R2380
 E  XX281
 E  L3581
 L  R3580
 E  XX581
 E  L5481
 L  R5480
 E  L5681
 L  R5680
 L  R1381
 L  R2381
 L  R3581
 L  R5481
 L  R5681
 E  HCP78
 E  HED78
 E  RES78
 E  RES79
 E  RES80
 E  RES81
 L  OUT2X
 E  HCP79
 E  HED79
 E  HCP80
 E  HED80
 L  CASCX
 E  HCP81
 E  HED81
 L  CHALX
 E  RCP78
 E  RCP79
 E  RCP80
 E  RCP81
 E  GCP78
 E  GCP79
 E  GCP80
 E  GCP81
 E  NCP78
 E  NCP79
 E  NCP80
 E  NCP81
 E  BCP78
 E  BCP79
 E  BCP80
 E  BCP81
 E  ACP78
 G  CRS78
 E  ACP79
 G  AAH78
 L  CCH78
 G  AAN78
 L  CCN78
 G  AAT78
 L  CCT78
 G  AAB78
 L  CCB78
 G  AAR78
 L  CCR78
 E  AAG78
 L  CCG78
 G  CRS79
 E  ACP80
 G  AAH79
 L  CCH79
 G  AAN79
 L  CCN79
 G  AAT79
 L  CCT79
 G  AAB79
 L  CCB79
 G  AAR79
 L  CCR79
 E  AAG79
 L  CCG79
 G  CRS80
 E  ACP81
 G  AAH80
 L  CCH80
 G  AAN80
 L  CCN80
 G  AAT80
 L  CCT80
 G  AAB80
 L  CCB80
 G  AAR80
 L  CCR80
 E  AAG80
 L  CCG80
 G  CRS81
 G  AAH81
 L  CCH81
 G  AAN81
 L  CCN81
 G  AAT81
 L  CCT81
 G  AAB81
 L  CCB81
 G  AAR81
 L  CCR81
 E  AAG81
 L  CCG81
 L  EEH78
 L  EEH79
 L  RPM78
 L  EEN78
 E  TCP78
 E  TCP79
 L  EET78
 L  EET79
 L  EET80
 L  EET81
 L  EEB78
 L  EEB79
 L  EEB80
 L  EEB81
 L  EER78
 L  EEG78
 L  EEH80
 L  RPM79
 L  EEN79
 E  TCP80
 L  EER79
 L  EEG79
 L  EEH81
 L  RPM80
 L  EEN80
 E  TCP81
 L  EER80
 L  EEG80
 L  RPM81
 L  EEN81
 L  EER81
 L  EEG81
 E  DGRES
COLUMNS
    VALRES    VLRES              -1.   OBJEC           .42409
    RNAI72    RAI72              -1.   CVI72               1.
    RNAI72    TRS72               1.   FPR72               1.
    RVAD72    RAI72               1.   OBJEC               1.
    DEPN72    DEP72              -1.   DEP73               1.
    DEPN72    RAI72              -1.   TRS72               1.
    INVT72    INV72              -1.   TRS72              -1.
    INTC72    INC72              -1.   FPR72               1.
    INTC72    TRS72               1.
    INTT72    INT72              -1.   CVI72         -1.25002
    INTT72    TRS72              -1.   FPR72              -1.
    DETT72    INT72             .085   TRS72               1.
    DETT72    TRS73              -1.   CAF72              -1.
    FOPR72    FPR72              -1.   FPR73               1.
    FOPR72    CAF72          4.00007
    RNAI73    RAI73              -1.   CVI73               1.
    RNAI73    TRS73               1.   FPR73               1.
    RVAD73    RAI73               1.   OBJEC           .90909
    DEPN73    DEP73              -1.   DEP74               1.
    DEPN73    RAI73              -1.   TRS73               1.
    INVT73    INV73              -1.   TRS73              -1.
    INTC73    INC73              -1.   FPR73               1.
    INTC73    TRS73               1.
    INTT73    INT73              -1.   CVI73         -1.25002
    INTT73    TRS73              -1.   FPR73              -1.
    DETT73    INT73             .085   TRS73               1.
    DETT73    TRS74              -1.   CAF73              -1.
    FOPR73    FPR73              -1.   FPR74               1.
    FOPR73    CAF73          4.00007
    RNAI74    RAI74              -1.   CVI74               1.
    RNAI74    TRS74               1.   FPR74               1.
    RVAD74    RAI74               1.   OBJEC           .82644
    DEPN74    DEP74              -1.   DEP75               1.
    DEPN74    RAI74              -1.   TRS74               1.
    INVT74    INV74              -1.   TRS74              -1.
    INTC74    INC74              -1.   FPR74               1.
    INTC74    TRS74               1.
    INTT74    INT74              -1.   CVI74         -1.25002
    INTT74    TRS74              -1.   FPR74              -1.
    DETT74    INT74             .085   TRS74               1.
    DETT74    TRS75              -1.   CAF74              -1.
    FOPR74    FPR74              -1.   FPR75               1.
    FOPR74    CAF74          4.00007
    RNAI75    RAI75              -1.   CVI75               1.
    RNAI75    TRS75               1.   FPR75               1.
    RVAD75    RAI75               1.   OBJEC           .75131
    DEPN75    DEP75              -1.   DEP76               1.
    DEPN75    RAI75              -1.   TRS75               1.
    INVT75    INV75              -1.   TRS75              -1.
    INTC75    INC75              -1.   FPR75               1.
    INTC75    TRS75               1.
    INTT75    INT75              -1.   CVI75         -1.25002
    INTT75    TRS75              -1.   FPR75              -1.
    DETT75    INT75             .085   TRS75               1.
    DETT75    TRS76              -1.   CAF75              -1.
    FOPR75    FPR75              -1.   FPR76               1.
    FOPR75    CAF75          4.00007
    RNAI76    RAI76              -1.   CVI76               1.
    RNAI76    TRS76               1.   FPR76               1.
    RVAD76    RAI76               1.   OBJEC           .68301
    DEPN76    DEP76              -1.   DEP77               1.
    DEPN76    RAI76              -1.   TRS76               1.
    INVT76    INV76              -1.   TRS76              -1.
    INTC76    INC76              -1.   FPR76               1.
    INTC76    TRS76               1.
    INTT76    INT76              -1.   CVI76         -1.25002
    INTT76    TRS76              -1.   FPR76              -1.
    DETT76    INT76             .085   TRS76               1.
    DETT76    TRS77              -1.   CAF76              -1.
    FOPR76    FPR76              -1.   FPR77               1.
    FOPR76    CAF76          4.00007
    RNAI77    RAI77              -1.   CVI77               1.
    RNAI77    TRS77               1.   FPR77               1.
    RVAD77    RAI77               1.   OBJEC           .62092
    DEPN77    DEP77              -1.   DEP78               1.
    DEPN77    RAI77              -1.   TRS77               1.
    INVT77    INV77              -1.   TRS77              -1.
    INTC77    INC77              -1.   FPR77               1.
    INTC77    TRS77               1.
    INTT77    INT77              -1.   CVI77         -1.25002
    INTT77    TRS77              -1.   FPR77              -1.
    DETT77    INT77             .085   TRS77               1.
    DETT77    TRS78              -1.   CAF77              -1.
    FOPR77    FPR77              -1.   FPR78               1.
    FOPR77    CAF77          4.00007
    RNAI78    RAI78              -1.   CVI78               1.
    RNAI78    TRS78               1.   FPR78               1.
    RVAD78    RAI78               1.   OBJEC           .56447
    DEPN78    DEP78              -1.   DEP79               1.
    DEPN78    RAI78              -1.   TRS78               1.
    INVT78    INV78              -1.   TRS78              -1.
    INTC78    INC78              -1.   FPR78               1.
    INTC78    TRS78               1.
    INTT78    INT78              -1.   CVI78         -1.25002
    INTT78    TRS78              -1.   FPR78              -1.
    DETT78    INT78             .085   TRS78               1.
    DETT78    TRS79              -1.   CAF78              -1.
    FOPR78    FPR78              -1.   FPR79               1.
    FOPR78    CAF78          4.00007
    RNAI79    RAI79              -1.   CVI79               1.
    RNAI79    TRS79               1.   FPR79               1.
    RVAD79    RAI79               1.   OBJEC           .51315
    DEPN79    DEP79              -1.   DEP80               1.
    DEPN79    RAI79              -1.   TRS79               1.
    INVT79    INV79              -1.   TRS79              -1.
    INTC79    INC79              -1.   FPR79               1.
    INTC79    TRS79               1.
    INTT79    INT79              -1.   CVI79         -1.25002
    INTT79    TRS79              -1.   FPR79              -1.
    DETT79    INT79             .085   TRS79               1.
    DETT79    TRS80              -1.   CAF79              -1.
    FOPR79    FPR79              -1.   FPR80               1.
    FOPR79    CAF79          4.00007
    RNAI80    RAI80              -1.   CVI80               1.
    RNAI80    TRS80               1.   FPR80               1.
    RVAD80    RAI80               1.   OBJEC            .4665
    DEPN80    DEP80              -1.   DEP81               1.
    DEPN80    RAI80              -1.   TRS80               1.
    INVT80    INV80              -1.   TRS80              -1.
    INTC80    INC80              -1.   FPR80               1.
    INTC80    TRS80               1.
    INTT80    INT80              -1.   CVI80         -1.25002
    INTT80    TRS80              -1.   FPR80              -1.
    DETT80    INT80             .085   TRS80               1.
    DETT80    TRS81              -1.   CAF80              -1.
    FOPR80    FPR80              -1.   FPR81               1.
    FOPR80    CAF80          4.00007
    RNAI81    RAI81              -1.   CVI81               1.
    RNAI81    TRS81               1.   FPR81               1.
    RVAD81    RAI81               1.   OBJEC           .42409
    DEPN81    DEP81              -1.   RAI81              -1.
    DEPN81    TRS81               1.
    INVT81    INV81              -1.   TRS81              -1.
    INTC81    INC81              -1.   FPR81               1.
    INTC81    TRS81               1.
    INTT81    INT81              -1.   CVI81         -1.25002
    INTT81    TRS81              -1.   FPR81              -1.
    INTT81    VLRES          9.99983
    DETT81    INT81             .085   TRS81               1.
    DETT81    CAF81              -1.
    FOPR81    FPR81              -1.   CAF81          4.00007
    CP1378    L1378              -1.   L1379               1.
    CP1378    R1378              -1.
    XX1378    R1378               1.   XX178              -1.
    XX1378    XX179               1.   XX378               1.
    XX1378    XX379              -1.
    CP2378    L2378              -1.   L2379               1.
    CP2378    R2378              -1.
    XX2378    R2378               1.   XX278              -1.
    XX2378    XX279               1.   XX378               1.
    XX2378    XX379              -1.
    CP3578    L3578              -1.   L3579               1.
    CP3578    R3578              -1.
    XX3578    R3578               1.   XX378              -1.
    XX3578    XX379               1.   XX578               1.
    XX3578    XX579              -1.
    CP5478    L5478              -1.   L5479               1.
    CP5478    R5478              -1.
    XX5478    R5478               1.   XX578              -1.
    XX5478    XX579               1.
    CP5678    L5678              -1.   L5679               1.
    CP5678    R5678              -1.
    XX5678    R5678               1.   XX578              -1.
    XX5678    XX579               1.
    CP1379    L1379              -1.   L1380               1.
    CP1379    R1379              -1.
    XX1379    R1379               1.   XX179              -1.
    XX1379    XX180               1.   XX379               1.
    XX1379    XX380              -1.
    CP2379    L2379              -1.   L2380               1.
    CP2379    R2379              -1.
    XX2379    R2379               1.   XX279              -1.
    XX2379    XX280               1.   XX379               1.
    XX2379    XX380              -1.
    CP3579    L3579              -1.   L3580               1.
    CP3579    R3579              -1.
    XX3579    R3579               1.   XX379              -1.
    XX3579    XX380               1.   XX579               1.
    XX3579    XX580              -1.
    CP5479    L5479              -1.   L5480               1.
    CP5479    R5479              -1.
    XX5479    R5479               1.   XX579              -1.
    XX5479    XX580               1.
    CP5679    L5679              -1.   L5680               1.
    CP5679    R5679              -1.
    XX5679    R5679               1.   XX579              -1.
    XX5679    XX580               1.
    CP1380    L1380              -1.   L1381               1.
    CP1380    R1380              -1.
    XX1380    R1380               1.   XX180              -1.
    XX1380    XX181               1.   XX380               1.
    XX1380    XX381              -1.
    CP2380    L2380              -1.   L2381               1.
    CP2380    R2380              -1.
    XX2380    R2380               1.   XX280              -1.
    XX2380    XX281               1.   XX380               1.
    XX2380    XX381              -1.
    CP3580    L3580              -1.   L3581               1.
    CP3580    R3580              -1.
    XX3580    R3580               1.   XX380              -1.
    XX3580    XX381               1.   XX580               1.
    XX3580    XX581              -1.
    CP5480    L5480              -1.   L5481               1.
    CP5480    R5480              -1.
    XX5480    R5480               1.   XX580              -1.
    XX5480    XX581               1.
    CP5680    L5680              -1.   L5681               1.
    CP5680    R5680              -1.
    XX5680    R5680               1.   XX580              -1.
    XX5680    XX581               1.
    CP1381    L1381              -1.   R1381              -1.
    XX1381    R1381               1.   XX181              -1.
    XX1381    XX381               1.
    CP2381    L2381              -1.   R2381              -1.
    XX2381    R2381               1.   XX281              -1.
    XX2381    XX381               1.
    CP3581    L3581              -1.   R3581              -1.
    XX3581    R3581               1.   XX381              -1.
    XX3581    XX581               1.
    CP5481    L5481              -1.   R5481              -1.
    XX5481    R5481               1.   XX581              -1.
    CP5681    L5681              -1.   R5681              -1.
    XX5681    R5681               1.   XX581              -1.
    FEXE78    FXE78              -1.   RAI78              -1.
    ACHT78    ACP78              -1.   CRS78           .11416
    ACHT78    ACP79               1.   AAH78              -1.
    ACHT78    CCH78          -.11416   AAN78              -1.
    ACHT78    CCN78          -.11416   AAT78              -1.
    ACHT78    CCT78          -.11416   AAB78              -1.
    ACHT78    CCB78          -.11416   AAR78              -1.
    ACHT78    CCR78          -.11416   AAG78              -1.
    ACHT78    CCG78          -.11416
    FEXE79    FXE79              -1.   RAI79              -1.
    ACHT79    ACP79              -1.   CRS79           .11416
    ACHT79    ACP80               1.   AAH79              -1.
    ACHT79    CCH79          -.11416   AAN79              -1.
    ACHT79    CCN79          -.11416   AAT79              -1.
    ACHT79    CCT79          -.11416   AAB79              -1.
    ACHT79    CCB79          -.11416   AAR79              -1.
    ACHT79    CCR79          -.11416   AAG79              -1.
    ACHT79    CCG79          -.11416
    FEXE80    FXE80              -1.   RAI80              -1.
    ACHT80    ACP80              -1.   CRS80           .11416
    ACHT80    ACP81               1.   AAH80              -1.
    ACHT80    CCH80          -.11416   AAN80              -1.
    ACHT80    CCN80          -.11416   AAT80              -1.
    ACHT80    CCT80          -.11416   AAB80              -1.
    ACHT80    CCB80          -.11416   AAR80              -1.
    ACHT80    CCR80          -.11416   AAG80              -1.
    ACHT80    CCG80          -.11416
    FEXE81    FXE81              -1.   RAI81              -1.
    ACHT81    ACP81              -1.   CRS81           .11416
    ACHT81    AAH81              -1.   CCH81          -.11416
    ACHT81    AAN81              -1.   CCN81          -.11416
    ACHT81    AAT81              -1.   CCT81          -.11416
    ACHT81    AAB81              -1.   CCB81          -.11416
    ACHT81    AAR81              -1.   CCR81          -.11416
    ACHT81    AAG81              -1.   CCG81          -.11416
    ESEC78    EEH78               1.   RAI78          -.35001
    ESEC78    OBJEC          -.28224
    GRES78    EEH78           .02433   EEH79          -.02508
    HEND78    HED78              -1.   HED79               1.
    HEND78    EEH78              -1.
    HCAP78    HCP78              -1.   HCP79               1.
    HCAP78    CRS78               1.   CCH78          -.98802
    HCAP78    CCN78          -.98802   CCT78          -.98802
    HCAP78    CCB78          -.98802   CCR78          -.98802
    HCAP78    CCG78          -.98802
    WK1H78    AAH78          8.76015   CCH78               1.
    WK2H78    AAH78           6.1321   CCH78               1.
    WK3H78    AAH78          4.02967   CCH78               1.
    WK4H78    AAH78          1.75203   CCH78               1.
    EHRP78    EEH78               1.   RPM78          -.75001
    EHRP78    RAI78          -.50001
    HENE78    RAI78          -.50001   AAH78              -1.
    HENE78    AAN78              -1.   AAT78              -1.
    HENE78    AAB78              -1.   AAR78              -1.
    HENE78    AAG78              -1.   EEH78               1.
    HENE78    OBJEC          -.28224
    NCAP78    NCP78              -1.   NCP79               1.
    NCAP78    CRS78               1.   EEN78         -7.00813
    NCAP78    CCN78          -.85001   CCT78          -.85001
    NCAP78    CCB78          -.85001   CCR78          -.85001
    NCAP78    CCG78          -.85001
    WK1N78    AAN78          8.76015   CCN78               1.
    WK2N78    AAN78           6.1321   CCN78               1.
    WK3N78    AAN78          4.02967   CCN78               1.
    WK4N78    AAN78          1.75203   CCN78               1.
    ENRP78    EEN78               1.   RPM78          -.75001
    ENRP78    RAI78          -.74282
    NENE78    RAI78          -.74282   AAN78              -1.
    NENE78    AAT78              -1.   AAB78              -1.
    NENE78    AAR78              -1.   AAG78              -1.
    NENE78    EEN78               1.
    TCAP78    TCP78              -1.   TCP79               1.
    TCAP78    CRS78               1.   EET78         -6.13201
    TCAP78    EET79            -.438   EET80           -.2628
    TCAP78    EET81           -.1752   CCT78             -.75
    TCAP78    CCT79             -.07   CCT80             -.03
    TCAP78    CCB78             -.75   CCB79             -.07
    TCAP78    CCB80             -.03   CCR78             -.75
    TCAP78    CCR79             -.07   CCR80             -.03
    TCAP78    CCG78             -.75   CCG79             -.07
    TCAP78    CCG80             -.03
    WK1T78    AAT78          8.76015   CCT78               1.
    WK2T78    AAT78           6.1321   CCT78               1.
    WK3T78    AAT78          4.02967   CCT78               1.
    WK4T78    AAT78          1.75203   CCT78               1.
    ETRP78    EET78               1.   RPM78          -.75001
    ETRP78    RAI78         -5.06437
    TENE78    RAI78         -5.06437   AAT78              -1.
    TENE78    AAB78              -1.   AAR78              -1.
    TENE78    AAG78              -1.   EET78               1.
    BCAP78    BCP78              -1.   BCP79               1.
    BCAP78    CRS78               1.   EEB78         -7.00813
    BCAP78    CCB78          -.75001   CCB79             -.05
    BCAP78    CCB80             -.03   CCB81             -.02
    BCAP78    CCR78          -.75001   CCR79             -.05
    BCAP78    CCR80             -.03   CCR81             -.02
    BCAP78    CCG78          -.75001   CCG79             -.05
    BCAP78    CCG80             -.03   CCG81             -.02
    WK1B78    AAB78          8.76015   CCB78               1.
    WK2B78    AAB78           6.1321   CCB78               1.
    WK3B78    AAB78          4.02967   CCB78               1.
    WK4B78    AAB78          1.75203   CCB78               1.
    EBRP78    EEB78               1.   RPM78          -.75001
    EBRP78    RAI78         -5.40147
    BENE78    RAI78         -5.40147   AAB78              -1.
    BENE78    AAR78              -1.   AAG78              -1.
    BENE78    EEB78               1.
    RCAP78    RCP78              -1.   RCP79               1.
    RCAP78    CRS78               1.   EER78         -2.19004
    RCAP78    CCR78          -.98802   CCG78          -.98802
    WK1R78    AAR78          8.76015   CCR78               1.
    WK2R78    AAR78           6.1321   CCR78               1.
    WK3R78    AAR78          4.02967   CCR78               1.
    WK4R78    AAR78          1.75203   CCR78               1.
    RENE78    AAR78              -1.   AAG78              -1.
    RENE78    EER78               1.   RPM78               1.
    GCAP78    GCP78              -1.   GCP79               1.
    GCAP78    CRS78               1.   EEG78         -3.50406
    GCAP78    CCG78          -.90002
    WK1G78    AAG78          8.76015   CCG78               1.
    WK2G78    AAG78           6.1321   CCG78               1.
    WK3G78    AAG78          4.02967   CCG78               1.
    WK4G78    AAG78          1.75203   CCG78               1.
    GENE78    RAI78        -16.09344   AAG78              -1.
    GENE78    EEG78               1.
    ESEC79    EEH79               1.   RAI79          -.35001
    ESEC79    OBJEC          -.25659
    GRES79    EEH79           .02433   EEH80          -.02508
    HEND79    HED79              -1.   HED80               1.
    HEND79    EEH79              -1.
    HCAP79    HCP79              -1.   HCP80               1.
    HCAP79    CRS79               1.   CCH79          -.98802
    HCAP79    CCN79          -.98802   CCT79          -.98802
    HCAP79    CCB79          -.98802   CCR79          -.98802
    HCAP79    CCG79          -.98802
    WK1H79    AAH79          8.76015   CCH79               1.
    WK2H79    AAH79           6.1321   CCH79               1.
    WK3H79    AAH79          4.02967   CCH79               1.
    WK4H79    AAH79          1.75203   CCH79               1.
    EHRP79    EEH79               1.   RPM79          -.75001
    EHRP79    RAI79          -.50001
    HENE79    RAI79          -.50001   AAH79              -1.
    HENE79    AAN79              -1.   AAT79              -1.
    HENE79    AAB79              -1.   AAR79              -1.
    HENE79    AAG79              -1.   EEH79               1.
    HENE79    OBJEC          -.25659
    NCAP79    NCP79              -1.   NCP80               1.
    NCAP79    CRS79               1.   EEN79         -7.00813
    NCAP79    CCN79          -.85001   CCT79          -.85001
    NCAP79    CCB79          -.85001   CCR79          -.85001
    NCAP79    CCG79          -.85001
    WK1N79    AAN79          8.76015   CCN79               1.
    WK2N79    AAN79           6.1321   CCN79               1.
    WK3N79    AAN79          4.02967   CCN79               1.
    WK4N79    AAN79          1.75203   CCN79               1.
    ENRP79    EEN79               1.   RPM79          -.75001
    ENRP79    RAI79          -.75025
    NENE79    RAI79          -.75025   AAN79              -1.
    NENE79    AAT79              -1.   AAB79              -1.
    NENE79    AAR79              -1.   AAG79              -1.
    NENE79    EEN79               1.
    TCAP79    TCP79              -1.   TCP80               1.
    TCAP79    CRS79               1.   EET79         -6.13201
    TCAP79    EET80            -.438   EET81           -.2628
    TCAP79    CCT79             -.75   CCT80             -.07
    TCAP79    CCT81             -.03   CCB79             -.75
    TCAP79    CCB80             -.07   CCB81             -.03
    TCAP79    CCR79             -.75   CCR80             -.07
    TCAP79    CCR81             -.03   CCG79             -.75
    TCAP79    CCG80             -.07   CCG81             -.03
    WK1T79    AAT79          8.76015   CCT79               1.
    WK2T79    AAT79           6.1321   CCT79               1.
    WK3T79    AAT79          4.02967   CCT79               1.
    WK4T79    AAT79          1.75203   CCT79               1.
    ETRP79    EET79               1.   RPM79          -.75001
    ETRP79    RAI79         -5.16566
    TENE79    RAI79         -5.16566   AAT79              -1.
    TENE79    AAB79              -1.   AAR79              -1.
    TENE79    AAG79              -1.   EET79               1.
    BCAP79    BCP79              -1.   BCP80               1.
    BCAP79    CRS79               1.   EEB79         -7.00813
    BCAP79    CCB79          -.75001   CCB80             -.05
    BCAP79    CCB81             -.03   CCR79          -.75001
    BCAP79    CCR80             -.05   CCR81             -.03
    BCAP79    CCG79          -.75001   CCG80             -.05
    BCAP79    CCG81             -.03
    WK1B79    AAB79          8.76015   CCB79               1.
    WK2B79    AAB79           6.1321   CCB79               1.
    WK3B79    AAB79          4.02967   CCB79               1.
    WK4B79    AAB79          1.75203   CCB79               1.
    EBRP79    EEB79               1.   RPM79          -.75001
    EBRP79    RAI79         -5.50949
    BENE79    RAI79         -5.50949   AAB79              -1.
    BENE79    AAR79              -1.   AAG79              -1.
    BENE79    EEB79               1.
    RCAP79    RCP79              -1.   RCP80               1.
    RCAP79    CRS79               1.   EER79         -2.19004
    RCAP79    CCR79          -.98802   CCG79          -.98802
    WK1R79    AAR79          8.76015   CCR79               1.
    WK2R79    AAR79           6.1321   CCR79               1.
    WK3R79    AAR79          4.02967   CCR79               1.
    WK4R79    AAR79          1.75203   CCR79               1.
    RENE79    AAR79              -1.   AAG79              -1.
    RENE79    EER79               1.   RPM79               1.
    GCAP79    GCP79              -1.   GCP80               1.
    GCAP79    CRS79               1.   EEG79         -3.50406
    GCAP79    CCG79          -.90002
    WK1G79    AAG79          8.76015   CCG79               1.
    WK2G79    AAG79           6.1321   CCG79               1.
    WK3G79    AAG79          4.02967   CCG79               1.
    WK4G79    AAG79          1.75203   CCG79               1.
    GENE79    RAI79         -16.4153   AAG79              -1.
    GENE79    EEG79               1.
    ESEC80    EEH80               1.   RAI80          -.35001
    ESEC80    OBJEC          -.23326
    GRES80    EEH80           .02433   EEH81          -.02508
    GRES80    DGRES              -1.
    HEND80    HED80              -1.   HED81               1.
    HEND80    EEH80              -1.
    HCAP80    HCP80              -1.   HCP81               1.
    HCAP80    CRS80               1.   CCH80          -.98802
    HCAP80    CCN80          -.98802   CCT80          -.98802
    HCAP80    CCB80          -.98802   CCR80          -.98802
    HCAP80    CCG80          -.98802
    WK1H80    AAH80          8.76015   CCH80               1.
    WK2H80    AAH80           6.1321   CCH80               1.
    WK3H80    AAH80          4.02967   CCH80               1.
    WK4H80    AAH80          1.75203   CCH80               1.
    EHRP80    EEH80               1.   RPM80          -.75001
    EHRP80    RAI80          -.50001
    HENE80    RAI80          -.50001   AAH80              -1.
    HENE80    AAN80              -1.   AAT80              -1.
    HENE80    AAB80              -1.   AAR80              -1.
    HENE80    AAG80              -1.   EEH80               1.
    HENE80    OBJEC          -.23326
    NCAP80    NCP80              -1.   NCP81               1.
    NCAP80    CRS80               1.   EEN80         -7.00813
    NCAP80    CCN80          -.85001   CCT80          -.85001
    NCAP80    CCB80          -.85001   CCR80          -.85001
    NCAP80    CCG80          -.85001
    WK1N80    AAN80          8.76015   CCN80               1.
    WK2N80    AAN80           6.1321   CCN80               1.
    WK3N80    AAN80          4.02967   CCN80               1.
    WK4N80    AAN80          1.75203   CCN80               1.
    ENRP80    EEN80               1.   RPM80          -.75001
    ENRP80    RAI80          -.75775
    NENE80    RAI80          -.75775   AAN80              -1.
    NENE80    AAT80              -1.   AAB80              -1.
    NENE80    AAR80              -1.   AAG80              -1.
    NENE80    EEN80               1.
    TCAP80    TCP80              -1.   TCP81               1.
    TCAP80    CRS80               1.   EET80         -6.13201
    TCAP80    EET81            -.438   CCT80             -.75
    TCAP80    CCT81             -.07   CCB80             -.75
    TCAP80    CCB81             -.07   CCR80             -.75
    TCAP80    CCR81             -.07   CCG80             -.75
    TCAP80    CCG81             -.07
    WK1T80    AAT80          8.76015   CCT80               1.
    WK2T80    AAT80           6.1321   CCT80               1.
    WK3T80    AAT80          4.02967   CCT80               1.
    WK4T80    AAT80          1.75203   CCT80               1.
    ETRP80    EET80               1.   RPM80          -.75001
    ETRP80    RAI80         -5.26896
    TENE80    RAI80         -5.26896   AAT80              -1.
    TENE80    AAB80              -1.   AAR80              -1.
    TENE80    AAG80              -1.   EET80               1.
    BCAP80    BCP80              -1.   BCP81               1.
    BCAP80    CRS80               1.   EEB80         -7.00813
    BCAP80    CCB80          -.75001   CCB81             -.05
    BCAP80    CCR80          -.75001   CCR81             -.05
    BCAP80    CCG80          -.75001   CCG81             -.05
    WK1B80    AAB80          8.76015   CCB80               1.
    WK2B80    AAB80           6.1321   CCB80               1.
    WK3B80    AAB80          4.02967   CCB80               1.
    WK4B80    AAB80          1.75203   CCB80               1.
    EBRP80    EEB80               1.   RPM80          -.75001
    EBRP80    RAI80         -5.61968
    BENE80    RAI80         -5.61968   AAB80              -1.
    BENE80    AAR80              -1.   AAG80              -1.
    BENE80    EEB80               1.
    RCAP80    RCP80              -1.   RCP81               1.
    RCAP80    CRS80               1.   EER80         -2.19004
    RCAP80    CCR80          -.98802   CCG80          -.98802
    WK1R80    AAR80          8.76015   CCR80               1.
    WK2R80    AAR80           6.1321   CCR80               1.
    WK3R80    AAR80          4.02967   CCR80               1.
    WK4R80    AAR80          1.75203   CCR80               1.
    RENE80    AAR80              -1.   AAG80              -1.
    RENE80    EER80               1.   RPM80               1.
    GCAP80    GCP80              -1.   GCP81               1.
    GCAP80    CRS80               1.   EEG80         -3.50406
    GCAP80    CCG80          -.90002
    WK1G80    AAG80          8.76015   CCG80               1.
    WK2G80    AAG80           6.1321   CCG80               1.
    WK3G80    AAG80          4.02967   CCG80               1.
    WK4G80    AAG80          1.75203   CCG80               1.
    GENE80    RAI80        -16.74359   AAG80              -1.
    GENE80    EEG80               1.
    ESEC81    EEH81               1.   RAI81          -.35001
    ESEC81    OBJEC          -.21205   VLRES         -1.50001
    GRES81    EEH81           .02433   DGRES               1.
    HEND81    HED81              -1.   EEH81              -1.
    HCAP81    HCP81              -1.   CRS81               1.
    HCAP81    CCH81          -.98802   CCN81          -.98802
    HCAP81    CCT81          -.98802   CCB81          -.98802
    HCAP81    CCR81          -.98802   CCG81          -.98802
    WK1H81    AAH81          8.76015   CCH81               1.
    WK2H81    AAH81           6.1321   CCH81               1.
    WK3H81    AAH81          4.02967   CCH81               1.
    WK4H81    AAH81          1.75203   CCH81               1.
    EHRP81    EEH81               1.   RPM81          -.75001
    EHRP81    RAI81          -.50001   VLRES          -.00009
    HENE81    RAI81          -.50001   AAH81              -1.
    HENE81    AAN81              -1.   AAT81              -1.
    HENE81    AAB81              -1.   AAR81              -1.
    HENE81    AAG81              -1.   EEH81               1.
    HENE81    OBJEC          -.21205   VLRES          -.00009
    NCAP81    NCP81              -1.   CRS81               1.
    NCAP81    EEN81         -7.00813   CCN81          -.85001
    NCAP81    CCT81          -.85001   CCB81          -.85001
    NCAP81    CCR81          -.85001   CCG81          -.85001
    WK1N81    AAN81          8.76015   CCN81               1.
    WK2N81    AAN81           6.1321   CCN81               1.
    WK3N81    AAN81          4.02967   CCN81               1.
    WK4N81    AAN81          1.75203   CCN81               1.
    ENRP81    EEN81               1.   RPM81          -.75001
    ENRP81    RAI81          -.76533   VLRES          8.58846
    NENE81    RAI81          -.76533   AAN81              -1.
    NENE81    AAT81              -1.   AAB81              -1.
    NENE81    AAR81              -1.   AAG81              -1.
    NENE81    EEN81               1.   VLRES          8.58846
    TCAP81    TCP81              -1.   CRS81               1.
    TCAP81    EET81         -6.13201   CCT81             -.75
    TCAP81    CCB81             -.75   CCR81             -.75
    TCAP81    CCG81             -.75
    WK1T81    AAT81          8.76015   CCT81               1.
    WK2T81    AAT81           6.1321   CCT81               1.
    WK3T81    AAT81          4.02967   CCT81               1.
    WK4T81    AAT81          1.75203   CCT81               1.
    ETRP81    EET81               1.   RPM81          -.75001
    ETRP81    RAI81         -5.37434   VLRES         68.52119
    TENE81    RAI81         -5.37434   AAT81              -1.
    TENE81    AAB81              -1.   AAR81              -1.
    TENE81    AAG81              -1.   EET81               1.
    TENE81    VLRES         68.52119
    BCAP81    BCP81              -1.   CRS81               1.
    BCAP81    EEB81         -7.00813   CCB81          -.75001
    BCAP81    CCR81          -.75001   CCG81          -.75001
    WK1B81    AAB81          8.76015   CCB81               1.
    WK2B81    AAB81           6.1321   CCB81               1.
    WK3B81    AAB81          4.02967   CCB81               1.
    WK4B81    AAB81          1.75203   CCB81               1.
    EBRP81    EEB81               1.   RPM81          -.75001
    EBRP81    RAI81         -5.73207   VLRES         73.08212
    BENE81    RAI81         -5.73207   AAB81              -1.
    BENE81    AAR81              -1.   AAG81              -1.
    BENE81    EEB81               1.   VLRES         73.08212
    RCAP81    RCP81              -1.   CRS81               1.
    RCAP81    EER81         -2.19004   CCR81          -.98802
    RCAP81    CCG81          -.98802
    WK1R81    AAR81          8.76015   CCR81               1.
    WK2R81    AAR81           6.1321   CCR81               1.
    WK3R81    AAR81          4.02967   CCR81               1.
    WK4R81    AAR81          1.75203   CCR81               1.
    RENE81    AAR81              -1.   AAG81              -1.
    RENE81    EER81               1.   RPM81               1.
    GCAP81    GCP81              -1.   CRS81               1.
    GCAP81    EEG81         -3.50406   CCG81          -.90002
    WK1G81    AAG81          8.76015   CCG81               1.
    WK2G81    AAG81           6.1321   CCG81               1.
    WK3G81    AAG81          4.02967   CCG81               1.
    WK4G81    AAG81          1.75203   CCG81               1.
    GENE81    RAI81        -17.07845   AAG81              -1.
    GENE81    EEG81               1.   VLRES        217.74481
    RESV78    RES78              -1.   CRS78              -1.
    RESV79    RES79              -1.   CRS79              -1.
    RESV80    RES80              -1.   CRS80              -1.
    RESV81    RES81              -1.   CRS81              -1.
    EGRP81    VLRES        217.74481
    AD1378    L1378           .50001   INV76          2.25538
    AD1378    INC76           .09195   INV77          1.57199
    AD1378    INC77           .24798   DEP78           .01688
    AD1378    FXE78           .23175   FXE79           .23754
    AD1378    FXE80           .24348   FXE81           .24957
    AD2378    L2378           .50001   INV76          3.38307
    AD2378    INC76           .13792   INV77          2.47302
    AD2378    INC77           .37666   DEP78           .02926
    AD2378    FXE78           .57937   FXE79           .59385
    AD2378    FXE80            .6087   FXE81           .62391
    AD3578    L3578          1.00002   INV76           .90215
    AD3578    INC76           .03678   INV77           1.1119
    AD3578    INC77           .11889   DEP78           .00788
    AD3578    FXE78           .08111   FXE79           .08314
    AD3578    FXE80           .08522   FXE81           .08735
    AD5478    L5478           .50001   INV76          2.25538
    AD5478    INC76           .09195   INV77          1.57199
    AD5478    INC77           .24798   DEP78           .01688
    AD5478    FXE78           .02317   FXE79           .02375
    AD5478    FXE80           .02435   FXE81           .02496
    AD5678    L5678               .2   INV76          5.63844
    AD5678    INC76           .22986   INV77           2.5497
    AD5678    INC77           .56367   DEP78           .06752
    AD5678    FXE78           .40556   FXE79           .41569
    AD5678    FXE80           .42609   FXE81           .43674
    AD1379    L1379           .50001   INV77          2.30048
    AD1379    INC77           .09378   INV78          1.60343
    AD1379    INC78           .25294   DEP79           .01722
    AD1379    FXE79           .23754   FXE80           .24348
    AD1379    FXE81           .24956
    AD2379    L2379           .50001   INV77          3.45072
    AD2379    INC77           .14068   INV78          2.52247
    AD2379    INC78           .38419   DEP79           .02985
    AD2379    FXE79           .59385   FXE80            .6087
    AD2379    FXE81           .62391
    AD3579    L3579          1.00002   INV77           .92019
    AD3579    INC77           .03751   INV78          1.13414
    AD3579    INC78           .12126   DEP79           .00804
    AD3579    FXE79           .08314   FXE80           .08522
    AD3579    FXE81           .08735
    AD5479    L5479           .50001   INV77          2.30048
    AD5479    INC77           .09378   INV78          1.60343
    AD5479    INC78           .25294   DEP79           .01722
    AD5479    FXE79           .02375   FXE80           .02435
    AD5479    FXE81           .02496
    AD5679    L5679               .2   INV77          5.75121
    AD5679    INC77           .23446   INV78          2.60069
    AD5679    INC78           .57495   DEP79           .06888
    AD5679    FXE79           .41569   FXE80           .42609
    AD5679    FXE81           .43674
    AD1380    L1380           .50001   INV78          2.34649
    AD1380    INC78           .09566   INV79           1.6355
    AD1380    INC79             .258   DEP80           .01756
    AD1380    FXE80           .24348   FXE81           .24956
    AD2380    L2380           .50001   INV78          3.51973
    AD2380    INC78           .14349   INV79          2.57292
    AD2380    INC79           .39187   DEP80           .03044
    AD2380    FXE80            .6087   FXE81           .62391
    AD3580    L3580          1.00002   INV78            .9386
    AD3580    INC78           .03826   INV79          1.15682
    AD3580    INC79           .12369   DEP80            .0082
    AD3580    FXE80           .08522   FXE81           .08735
    AD5480    L5480           .50001   INV78          2.34649
    AD5480    INC78           .09566   INV79           1.6355
    AD5480    INC79             .258   DEP80           .01756
    AD5480    FXE80           .02435   FXE81           .02496
    AD5680    L5680               .2   INV78          5.86623
    AD5680    INC78           .23915   INV79           2.6527
    AD5680    INC79           .58645   DEP80           .07025
    AD5680    FXE80           .42609   FXE81           .43674
    AD1381    L1381           .50001   INV79          2.39342
    AD1381    INC79           .09757   INV80          1.66821
    AD1381    INC80           .26316   DEP81           .01791
    AD1381    FXE81           .24956
    AD2381    L2381           .50001   INV79          3.59013
    AD2381    INC79           .14636   INV80          2.62438
    AD2381    INC80           .39971   DEP81           .03105
    AD2381    FXE81           .62391
    AD3581    L3581          1.00002   INV79           .95737
    AD3581    INC79           .03903   INV80          1.17995
    AD3581    INC80           .12616   DEP81           .00836
    AD3581    FXE81           .08735
    AD5481    L5481           .50001   INV79          2.39342
    AD5481    INC79           .09757   INV80          1.66821
    AD5481    INC80           .26316   DEP81           .01791
    AD5481    FXE81           .02496
    AD5681    L5681               .2   INV79          5.98354
    AD5681    INC79           .24393   INV80          2.70575
    AD5681    INC80           .59817   DEP81           .07166
    AD5681    FXE81           .43674
    OUT278    INV72          1.65574   INC72           .06486
    OUT278    INV73          5.67333   INC73           .34686
    OUT278    INV74         22.61821   INC74           1.4279
    OUT278    INV75         47.27449   INC75          4.05381
    OUT278    INV76         53.49582   INC76          7.68351
    OUT278    INV77         32.85997   INC77         10.46422
    OUT278    DEP78          1.43546   FXE78           .50139
    OUT278    FXE79           .52577   FXE80           .55136
    OUT278    FXE81           .57821   HCP78           .39701
    OUT278    HED78          2.07003   RES78           .02779
    OUT278    RES79           .02779   RES80           .02724
    OUT278    RES81           .01264   VLRES         27.45341
    OUT278    OUT2X               1.
    OUT279    INV73          1.72197   INC73           .06745
    OUT279    INV74          5.90026   INC74           .36074
    OUT279    INV75         23.52068   INC75          1.48493
    OUT279    INV76         49.16049   INC76           4.2156
    OUT279    INV77         55.61571   INC77          7.98954
    OUT279    INV78         34.14847   INC78         10.87979
    OUT279    DEP79          1.49269   FXE79           .52577
    OUT279    FXE80           .55136   FXE81           .57821
    OUT279    HCP79           .39701   HED79          2.07003
    OUT279    RES79           .02779   RES80           .02751
    OUT279    RES81           .02696   VLRES         27.96248
    OUT279    OUT2X               1.
    OUT280    INV74          1.79085   INC74           .07015
    OUT280    INV75          6.13627   INC75           .37517
    OUT280    INV76         24.45918   INC76          1.54423
    OUT280    INV77          51.1218   INC77          4.38385
    OUT280    INV78         57.81995   INC78          8.30778
    OUT280    INV79         35.48801   INC79          11.3119
    OUT280    DEP80           1.5522   FXE80           .55136
    OUT280    FXE81           .57821   HCP80           .39701
    OUT280    HED80          2.07003   RES80           .02751
    OUT280    RES81           .02724   VLRES         28.49783
    OUT280    OUT2X               1.
    CASC78    INV74          3.50541   INC74           .13731
    CASC78    INV75         13.65878   INC75           .79889
    CASC78    INV76         18.45009   INC76          1.99405
    CASC78    INV77         13.03732   INC77          3.07123
    CASC78    DEP78           .44353   FXE78           .30957
    CASC78    FXE79           .32504   FXE80            .3413
    CASC78    FXE81           .35836   HCP78             .224
    CASC78    HED78           .46701   RES78           .01568
    CASC78    RES79           .01568   RES80           .01537
    CASC78    RES81           .00713   XX378             .224
    CASC78    VLRES         12.32012   CASCX               1.
    CASC79    INV75          3.64563   INC75            .1428
    CASC79    INV76         14.20513   INC76           .83085
    CASC79    INV77          19.1881   INC77          2.07381
    CASC79    INV78         13.55881   INC78          3.19407
    CASC79    DEP79           .46127   FXE79           .32504
    CASC79    FXE80            .3413   FXE81           .35836
    CASC79    HCP79             .224   HED79           .46701
    CASC79    RES79           .01568   RES80           .01552
    CASC79    RES81           .01521   XX379             .224
    CASC79    VLRES         12.47792   CASCX               1.
    CASC80    INV76          3.79145   INC76           .14852
    CASC80    INV77         14.77332   INC77           .86408
    CASC80    INV78          19.9556   INC78          2.15676
    CASC80    INV79         14.10115   INC79          3.32183
    CASC80    DEP80           .47972   FXE80            .3413
    CASC80    FXE81           .35836   HCP80             .224
    CASC80    HED80           .46701   RES80           .01552
    CASC80    RES81           .01537   XX380             .224
    CASC80    VLRES         12.64388   CASCX               1.
    CASC81    INV77           3.9431   INC77           .15446
    CASC81    INV78         15.36425   INC78           .89865
    CASC81    INV79         20.75381   INC79          2.24303
    CASC81    INV80         14.66519   INC80          3.45471
    CASC81    DEP81           .49891   FXE81           .35836
    CASC81    HCP81             .224   HED81           .46701
    CASC81    RES81           .01537   XX381             .224
    CASC81    VLRES         12.81823   CASCX               1.
    CHAL78    INV74           2.9618   INC74           .11602
    CHAL78    INV75         13.08475   INC75           .73549
    CHAL78    INV76         17.00111   INC76          1.85637
    CHAL78    INV77         12.92585   INC77          2.88321
    CHAL78    DEP78            .3985   FXE78           .34173
    CHAL78    FXE79           .35785   FXE80           .37476
    CHAL78    FXE81           .39249   HCP78             .124
    CHAL78    HED78              .19   RES78           .00868
    CHAL78    RES79           .00868   RES80           .00851
    CHAL78    RES81           .00395   XX378             .124
    CHAL78    VLRES         12.26072   CHALX               1.
    CHAL79    INV75          3.07718   INC75           .12054
    CHAL79    INV76          13.6013   INC76           .76441
    CHAL79    INV77         17.65373   INC77          1.92882
    CHAL79    INV78         13.40728   INC78           2.9944
    CHAL79    DEP79           .41417   FXE79           .35785
    CHAL79    FXE80           .37476   FXE81           .39249
    CHAL79    HCP79             .124   HED79              .19
    CHAL79    RES79           .00868   RES80           .00859
    CHAL79    RES81           .00842   XX379             .124
    CHAL79    VLRES         12.40026   CHALX               1.
    CHAL80    INV76          3.19711   INC76           .12523
    CHAL80    INV77         14.13837   INC77           .79447
    CHAL80    INV78         18.33189   INC78          2.00413
    CHAL80    INV79         13.90725   INC79          3.10996
    CHAL80    DEP80           .43047   FXE80           .37476
    CHAL80    FXE81           .39249   HCP80             .124
    CHAL80    HED80              .19   RES80           .00859
    CHAL80    RES81           .00851   XX380             .124
    CHAL80    VLRES         12.54695   CHALX               1.
    CHAL81    INV77          3.32177   INC77           .13012
    CHAL81    INV78         14.69678   INC78           .82573
    CHAL81    INV79         19.03664   INC79          2.08242
    CHAL81    INV80          14.4265   INC80          3.23006
    CHAL81    DEP81           .44742   FXE81           .39249
    CHAL81    HCP81             .124   HED81              .19
    CHAL81    RES81           .00851   XX381             .124
    CHAL81    VLRES         12.70101   CHALX               1.
    CHAI78    INV72          1.15012   INC72           .04505
    CHAI78    INV73         11.88958   INC73            .5523
    CHAI78    INV74         26.33287   INC74          2.00625
    CHAI78    INV75          38.2083   INC75          4.37723
    CHAI78    INV76         55.27832   INC76          7.69628
    CHAI78    INV77         36.20004   INC77         10.67664
    CHAI78    DEP78           1.4172   FXE78           .61043
    CHAI78    FXE79           .63754   FXE80           .66594
    CHAI78    FXE81           .69569   RCP78           .80001
    CHAI78    RES78             .088   RES79           .08554
    CHAI78    RES80           .08311   RES81           .06272
    CHAI78    XX378           .80001   VLRES         28.90616
    CHAI79    INV73          1.19612   INC73           .04685
    CHAI79    INV74         12.36516   INC74           .57439
    CHAI79    INV75         27.37502   INC75          2.08606
    CHAI79    INV76         39.71201   INC76          4.55051
    CHAI79    INV77         57.39073   INC77          7.99763
    CHAI79    INV78          37.5199   INC78         11.08884
    CHAI79    DEP79          1.47295   FXE79           .63754
    CHAI79    FXE80           .66594   FXE81           .69569
    CHAI79    RCP79           .80001   RES79           .08712
    CHAI79    RES80           .08467   RES81           .08311
    CHAI79    XX379           .80001   VLRES          29.4023
    CHAI80    INV74          1.24397   INC74           .04873
    CHAI80    INV75         12.85976   INC75           .59737
    CHAI80    INV76          28.4586   INC76          2.16906
    CHAI80    INV77         41.27533   INC77          4.73069
    CHAI80    INV78          59.5856   INC78          8.31091
    CHAI80    INV79            38.89   INC79         11.51721
    CHAI80    DEP80          1.53091   FXE80           .66594
    CHAI80    FXE81           .69569   RCP80           .80001
    CHAI80    RES80           .08624   RES81           .08467
    CHAI80    XX380           .80001   VLRES         29.92389
    CHAI81    INV75          1.29373   INC75           .05068
    CHAI81    INV76         13.37415   INC76           .62126
    CHAI81    INV77          29.5853   INC77          2.25536
    CHAI81    INV78         42.90071   INC78          4.91803
    CHAI81    INV79         61.86629   INC79          8.63658
    CHAI81    INV80         40.31227   INC80         11.96242
    CHAI81    DEP81          1.59116   FXE81           .69569
    CHAI81    RCP81           .80001   RES81           .08624
    CHAI81    XX381           .80001   VLRES         30.47165
    TGAZ78    INV76           .88587   INC76            .0347
    TGAZ78    INV77          5.88239   INC77            .2971
    TGAZ78    DEP78           .26102   FXE78           .06932
    TGAZ78    FXE79           .07265   FXE80           .07614
    TGAZ78    FXE81           .07981   GCP78              .06
    TGAZ78    RES78            .0075   RES79           .00731
    TGAZ78    RES80           .00719   RES81           .00706
    TGAZ78    XX378              .06   VLRES           5.0256
    TGAZ79    INV77           .91423   INC77           .03581
    TGAZ79    INV78          6.08514   INC78           .30718
    TGAZ79    DEP79              .27   FXE79           .07265
    TGAZ79    FXE80           .07614   FXE81           .07981
    TGAZ79    GCP79              .06   RES79           .00743
    TGAZ79    RES80           .00731   RES81           .00711
    TGAZ79    XX379              .06   VLRES          5.06904
    TGAZ80    INV78           .94353   INC78           .03696
    TGAZ80    INV79          6.29493   INC79            .3176
    TGAZ80    DEP80           .27928   FXE80           .07614
    TGAZ80    FXE81           .07981   GCP80              .06
    TGAZ80    RES80           .00743   RES81           .00723
    TGAZ80    XX380              .06   VLRES          5.11833
    TGAZ81    INV79            .9738   INC79           .03814
    TGAZ81    INV80          6.51199   INC80           .32838
    TGAZ81    DEP81           .28888   FXE81           .07981
    TGAZ81    GCP81              .06   RES81           .00735
    TGAZ81    XX381              .06   VLRES          5.17342
    N60078    INV72         16.92413   INC72           .66294
    N60078    INV73         36.58653   INC73          2.70708
    N60078    INV74         65.80643   INC74          6.50584
    N60078    INV75          77.6349   INC75         11.61492
    N60078    INV76         76.10367   INC76         16.72708
    N60078    INV77          64.4689   INC77         20.92299
    N60078    DEP78          7.19542   FXE78          3.01304
    N60078    FXE79           3.0733   FXE80          3.13476
    N60078    FXE81          3.19745   NCP78           .60001
    N60078    RES78           .29101   RES79            .2287
    N60078    RES80           .16759   RES81           .15133
    N60078    XX178           .60001   VLRES         129.2112
    N60079    INV73         17.60109   INC73           .68946
    N60079    INV74         38.04999   INC74          2.81536
    N60079    INV75         68.43272   INC75          6.76584
    N60079    INV76         80.72719   INC76         12.07855
    N60079    INV77         79.09523   INC77         17.39268
    N60079    INV78         66.97932   INC78         21.75198
    N60079    DEP79          7.48138   FXE79           3.0733
    N60079    FXE80          3.13476   FXE81          3.19745
    N60079    NCP79           .60001   RES79            .2881
    N60079    RES80           .22639   RES81           .16588
    N60079    XX179           .60001   VLRES        131.16989
    N60080    INV74         18.30511   INC74           .71703
    N60080    INV75         39.57196   INC75          2.92797
    N60080    INV76         71.16388   INC76          7.03623
    N60080    INV77         83.94284   INC77         12.56071
    N60080    INV78          82.2054   INC78         18.08487
    N60080    INV79         69.58878   INC79         22.61397
    N60080    DEP80          7.77874   FXE80          3.13476
    N60080    FXE81          3.19745   NCP80           .60001
    N60080    RES80           .28519   RES81           .22408
    N60080    XX180           .60001   VLRES        133.29187
    N60081    INV75         19.03732   INC75           .74571
    N60081    INV76         41.15482   INC76          3.04509
    N60081    INV77         74.00421   INC77          7.31744
    N60081    INV78         87.28688   INC78         13.06213
    N60081    INV79         85.43887   INC79         18.80464
    N60081    INV80         72.30124   INC80         23.51027
    N60081    DEP81          8.08795   FXE81          3.19745
    N60081    NCP81           .60001   RES81           .28228
    N60081    XX181           .60001   VLRES        135.57887
    T75078    INV73          2.25458   INC73           .08831
    T75078    INV74          16.5063   INC74           .81628
    T75078    INV75         31.35274   INC75          2.62702
    T75078    INV76         60.29823   INC76          6.01129
    T75078    INV77         38.60837   INC77          9.41463
    T75078    DEP78           3.3623   FXE78          1.76342
    T75078    FXE79          1.80724   FXE80          1.85216
    T75078    FXE81          1.89819   BCP78           .75001
    T75078    RES78           .43126   RES79            .3527
    T75078    RES80           .27563   RES81           .23645
    T75078    XX278           .75001   VLRES         67.22089
    T75079    INV74          2.34476   INC74           .09185
    T75079    INV75         17.16382   INC75           .84883
    T75079    INV76         32.60081   INC76          2.73166
    T75079    INV77         62.68602   INC77          6.25015
    T75079    INV78         40.12126   INC78          9.78758
    T75079    DEP79          3.49594   FXE79          1.80724
    T75079    FXE80          1.85216   FXE81          1.89819
    T75079    BCP79           .75001   RES79           .42695
    T75079    RES80           .34914   RES81           .27282
    T75079    XX279           .75001   VLRES         68.13623
    T75080    INV75          2.43855   INC75           .09552
    T75080    INV76         17.84756   INC76           .88267
    T75080    INV77         33.89868   INC77          2.84048
    T75080    INV78         65.16879   INC78          6.49853
    T75080    INV79         41.69402   INC79         10.17536
    T75080    DEP80           3.6349   FXE80          1.85216
    T75080    FXE81          1.89819   BCP80           .75001
    T75080    RES80           .42264   RES81           .34557
    T75080    XX280           .75001   VLRES         69.12788
    T75081    INV76           2.5361   INC76           .09934
    T75081    INV77         18.55862   INC77           .91786
    T75081    INV78         35.24835   INC78          2.95364
    T75081    INV79         67.75043   INC79          6.75682
    T75081    INV80         43.32906   INC80         10.57858
    T75081    DEP81          3.77941   FXE81          1.89819
    T75081    BCP81           .75001   RES81           .41833
    T75081    XX281           .75001   VLRES         70.19667
RHS
    RHS1      RAI72       -329.00562   TRS72         73.00125
    RHS1      FPR72       -1048.0178   DEP72        -65.00111
    RHS1      DEP73         -5.00011   INV72       -349.00586
    RHS1      INC72        -54.00093   INT72       -192.00334
    RHS1      TRS73        139.82237   CAF72        2846.0491
    RHS1      FPR73          -6.0001   RAI73       -336.45654
    RHS1      DEP74         -4.00006   INV73       -313.00537
    RHS1      INC73        -50.00087   INT73        -191.0033
    RHS1      TRS74        148.63713   CAF73        2697.0461
    RHS1      FPR74          -6.0001   RAI74        -343.0564
    RHS1      DEP75         -4.00006   INV74       -302.00513
    RHS1      INC74        -49.00082   INT74        -184.0032
    RHS1      TRS75         160.4463   CAF74        2539.0432
    RHS1      FPR75         -5.00009   RAI75       -345.83228
    RHS1      DEP76         -9.00015   INV75       -320.00562
    RHS1      INC75        -60.00102   INT75       -174.00296
    RHS1      TRS76        100.24867   CAF75          2369.04
    RHS1      FPR76         -4.00007   RAI76       -367.80737
    RHS1      DEP77         -4.00006   INV76       -267.00464
    RHS1      INC76        -65.00111   INT76       -167.00285
    RHS1      TRS77        105.04611   CAF76        2259.0381
    RHS1      FPR77         -4.00007   RAI77       -387.00073
    RHS1      DEP78         -3.00003   INV77       -231.00404
    RHS1      INC77        -33.00056   INT77       -161.00279
    RHS1      TRS78        117.83769   CAF77        2144.0359
    RHS1      FPR78         -3.00005   RAI78       -399.19189
    RHS1      DEP79         -5.00009   INV78       -190.00325
    RHS1      INC78        -30.00052   INT78        -154.0027
    RHS1      TRS79         98.62241   CAF78        2016.0334
    RHS1      FPR79         -2.00004   RAI79       -427.51025
    RHS1      DEP80         -3.00011   INV79       -183.00316
    RHS1      INC79        -31.00053   INT79       -151.00256
    RHS1      TRS80        113.40132   CAF79        1907.0313
    RHS1      FPR80         -1.00002   RAI80       -456.94946
    RHS1      DEP81         -7.00009   INV80       -192.00334
    RHS1      INC80        -30.00052   INT80       -144.00246
    RHS1      TRS81        161.17412   CAF80        1783.0288
    RHS1      FPR81         -1.00002   RAI81       -487.49097
    RHS1      INV81       -202.00348   INC81        -30.00052
    RHS1      INT81       -137.00235   CAF81        1611.0256
    RHS1      L1378         -3.00005   XX178         -2.99005
    RHS1      XX179             -.11   XX378               .1
    RHS1      XX379              -.1   L2378         -4.00007
    RHS1      L2379              -.1   XX278         -3.90007
    RHS1      XX279               .2   L3578         -7.00012
    RHS1      XX578           .70001   L5478         -4.00007
    RHS1      L5678         -3.00005   L5679           .90002
    RHS1      XX180          -.40001   XX280               .2
    RHS1      XX580              -.1   XX281               .1
    RHS1      HCP78        -10.00017   HED78        -93.46956
    RHS1      RES78         -2.30004   RES79         -2.20004
    RHS1      RES80         -2.10003   RES81         -2.10003
    RHS1      OUT2X               1.   CASCX               1.
    RHS1      CHALX               1.   RCP78              -4.
    RHS1      GCP78          -.50001   NCP78         -1.00002
    RHS1      ACP78          -.35501   CRS78         15.35027
    RHS1      ACP79             .037   CCT78             .085
    RHS1      CCB78             .085   CCR78             .085
    RHS1      CCG78             .085   CRS79         16.56027
    RHS1      CCT79              .03   CCB79              .03
    RHS1      CCR79              .03   CCG79              .03
    RHS1      CRS80         17.86028   ACP81             .202
    RHS1      CRS81         19.27029   EEH78          7.12154
    RHS1      TCP78         -1.00002   EET78           .62196
    RHS1      EET79            .3504   EET80            .1752
BOUNDS
 FR BNDS1     RVAD72
 FR BNDS1     RVAD73
 FR BNDS1     RVAD74
 FR BNDS1     RVAD75
 FR BNDS1     RVAD76
 FR BNDS1     RVAD77
 FR BNDS1     RVAD78
 FR BNDS1     RVAD79
 FR BNDS1     RVAD80
 FR BNDS1     RVAD81
 FR BNDS1     ACHT78
 FR BNDS1     ACHT79
 FR BNDS1     ACHT80
 FR BNDS1     ACHT81
 UP BNDS1     GRES78       284.00488
 UP BNDS1     WK1H78         8.07907
 UP BNDS1     WK2H78         1.79244
 UP BNDS1     WK3H78         1.57957
 UP BNDS1     WK4H78         1.51985
 UP BNDS1     WK1N78         8.07907
 UP BNDS1     WK2N78         1.79244
 UP BNDS1     WK3N78         1.57957
 UP BNDS1     WK4N78         1.51985
 UP BNDS1     WK1T78         8.07907
 UP BNDS1     WK2T78         1.79244
 UP BNDS1     WK3T78         1.57957
 UP BNDS1     WK4T78         1.51985
 UP BNDS1     WK1B78         8.07907
 UP BNDS1     WK2B78         1.79244
 UP BNDS1     WK3B78         1.57957
 UP BNDS1     WK4B78         1.51985
 UP BNDS1     WK1R78         8.07907
 UP BNDS1     WK2R78         1.79244
 UP BNDS1     WK3R78         1.57957
 UP BNDS1     WK4R78         1.51985
 FX BNDS1     WK1G78         8.07907
 FX BNDS1     WK2G78         1.79244
 FX BNDS1     WK3G78         1.57957
 FX BNDS1     WK4G78         1.51985
 UP BNDS1     GRES79       284.00488
 UP BNDS1     WK1H79         8.71593
 UP BNDS1     WK2H79         1.93373
 UP BNDS1     WK3H79         1.70408
 UP BNDS1     WK4H79         1.63966
 UP BNDS1     WK1N79         8.71593
 UP BNDS1     WK2N79         1.93373
 UP BNDS1     WK3N79         1.70408
 UP BNDS1     WK4N79         1.63966
 UP BNDS1     WK1T79         8.71593
 UP BNDS1     WK2T79         1.93373
 UP BNDS1     WK3T79         1.70408
 UP BNDS1     WK4T79         1.63966
 UP BNDS1     WK1B79         8.71593
 UP BNDS1     WK2B79         1.93373
 UP BNDS1     WK3B79         1.70408
 UP BNDS1     WK4B79         1.63966
 UP BNDS1     WK1R79         8.71593
 UP BNDS1     WK2R79         1.93373
 UP BNDS1     WK3R79         1.70408
 UP BNDS1     WK4R79         1.63966
 FX BNDS1     WK1G79         8.71593
 FX BNDS1     WK2G79         1.93373
 FX BNDS1     WK3G79         1.70408
 FX BNDS1     WK4G79         1.63966
 UP BNDS1     GRES80       284.00488
 UP BNDS1     WK1H80         9.40015
 UP BNDS1     WK2H80         2.08553
 UP BNDS1     WK3H80         1.83785
 UP BNDS1     WK4H80         1.76837
 UP BNDS1     WK1N80         9.40015
 UP BNDS1     WK2N80         2.08553
 UP BNDS1     WK3N80         1.83785
 UP BNDS1     WK4N80         1.76837
 UP BNDS1     WK1T80         9.40015
 UP BNDS1     WK2T80         2.08553
 UP BNDS1     WK3T80         1.83785
 UP BNDS1     WK4T80         1.76837
 UP BNDS1     WK1B80         9.40015
 UP BNDS1     WK2B80         2.08553
 UP BNDS1     WK3B80         1.83785
 UP BNDS1     WK4B80         1.76837
 UP BNDS1     WK1R80         9.40015
 UP BNDS1     WK2R80         2.08553
 UP BNDS1     WK3R80         1.83785
 UP BNDS1     WK4R80         1.76837
 FX BNDS1     WK1G80         9.40015
 FX BNDS1     WK2G80         2.08553
 FX BNDS1     WK3G80         1.83785
 FX BNDS1     WK4G80         1.76837
 UP BNDS1     GRES81       284.00488
 UP BNDS1     WK1H81        10.14226
 UP BNDS1     WK2H81         2.25018
 UP BNDS1     WK3H81         1.98295
 UP BNDS1     WK4H81         1.90798
 UP BNDS1     WK1N81        10.14226
 UP BNDS1     WK2N81         2.25018
 UP BNDS1     WK3N81         1.98295
 UP BNDS1     WK4N81         1.90798
 UP BNDS1     WK1T81        10.14226
 UP BNDS1     WK2T81         2.25018
 UP BNDS1     WK3T81         1.98295
 UP BNDS1     WK4T81         1.90798
 UP BNDS1     WK1B81        10.14226
 UP BNDS1     WK2B81         2.25018
 UP BNDS1     WK3B81         1.98295
 UP BNDS1     WK4B81         1.90798
 UP BNDS1     WK1R81        10.14226
 UP BNDS1     WK2R81         2.25018
 UP BNDS1     WK3R81         1.98295
 UP BNDS1     WK4R81         1.90798
 FX BNDS1     WK1G81        10.14226
 FX BNDS1     WK2G81         2.25018
 FX BNDS1     WK3G81         1.98295
 FX BNDS1     WK4G81         1.90798
 UP BNDS1     AD1378             10.
 UP BNDS1     AD2378             10.
 UP BNDS1     AD3578             10.
 UP BNDS1     AD5478             10.
 UP BNDS1     AD5678             10.
 UP BNDS1     AD1379             10.
 UP BNDS1     AD2379             10.
 UP BNDS1     AD3579             10.
 UP BNDS1     AD5479             10.
 UP BNDS1     AD5679             10.
 UP BNDS1     AD1380             10.
 UP BNDS1     AD2380             10.
 UP BNDS1     AD3580             10.
 UP BNDS1     AD5480             10.
 UP BNDS1     AD5680             10.
 UP BNDS1     AD1381             10.
 UP BNDS1     AD2381             10.
 UP BNDS1     AD3581             10.
 UP BNDS1     AD5481             10.
 UP BNDS1     AD5681             10.
 UP BNDS1     OUT278              1.
 UP BNDS1     OUT279              1.
 UP BNDS1     OUT280              1.
 UP BNDS1     CASC78              1.
 UP BNDS1     CASC79              1.
 UP BNDS1     CASC80              1.
 UP BNDS1     CASC81              1.
 UP BNDS1     CHAL78              1.
 UP BNDS1     CHAL79              1.
 UP BNDS1     CHAL80              1.
 UP BNDS1     CHAL81              1.
 UP BNDS1     CHAI78              1.
 UP BNDS1     CHAI79              1.
 UP BNDS1     CHAI80              1.
 UP BNDS1     CHAI81              1.
 UP BNDS1     TGAZ78             10.
 UP BNDS1     TGAZ79             10.
 UP BNDS1     TGAZ80             10.
 UP BNDS1     TGAZ81             10.
 UP BNDS1     N60078              2.
 UP BNDS1     N60079              2.
 UP BNDS1     N60080              2.
 UP BNDS1     N60081              2.
 UP BNDS1     T75078              2.
 UP BNDS1     T75079              2.
 UP BNDS1     T75080              2.
 UP BNDS1     T75081              2.
ENDATA
