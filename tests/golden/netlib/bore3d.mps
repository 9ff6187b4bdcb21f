NAME          BORE3D
ROWS
 E  B...XI
 E  TIE.MRAR
 E  BAB...XI
 E  BAC...XI
 E  BAD...XI
 E  BAF...XI
 E  BAG...XI
 E  BAH...XI
 E  BAN...XI
 E  BAR...XI
 E  BAS...XI
 E  BAY...XI
 E  BBT...XI
 E  BCC...XI
 E  BCE...XI
 E  BCF...XI
 E  BCK...XI
 E  BC2...XI
 E  BC3...XI
 E  BC4...XI
 E  BDC...XI
 E  BDF...XI
 E  BDH...XI
 E  BD1...XI
 E  BFF...XI
 E  BFG...XI
 E  BFH...XI
 E  BFI...XI
 E  BFL...XI
 E  BFN...XI
 E  BFP...XI
 E  BFS...XI
 E  BFX...XI
 E  BF1...XI
 E  BF2...XI
 E  BF3...XI
 E  BF4...XI
 E  BGF...XI
 E  BG1...XI
 E  BHB...XI
 E  BHF...XI
 E  BHG...XI
 E  BHH...XI
 E  BHN...XI
 E  BHP...XI
 E  BHR...XI
 E  BHS...XI
 E  BH1...XI
 E  BIC...XI
 E  BIF...XI
 E  BIL...XI
 E  BIP...XI
 E  BIS...XI
 E  BL2...XI
 E  BL3...XI
 E  BL4...XI
 E  BMM...XI
 E  BMR...XI
 E  BNF...XI
 E  BNH...XI
 E  BNP...XI
 E  BN1...XI
 E  BO2...XI
 E  BPR...XI
 E  BPS...XI
 E  BPX...XI
 E  BP1...XI
 E  BRS...XI
 E  BSB...XI
 E  BSS...XI
 E  BWP...XI
 E  BWS...XI
 E  BWW...XI
 E  BYD...XI
 E  BYF...XI
 E  BYG...XI
 E  BYH...XI
 E  BYN...XI
 E  BYS...XI
 E  BYY...XI
 E  B6H...XI
 E  B9H...XI
 E  BFH.VOXI
 E  BFL.VOXI
 E  BFN.VOXI
 E  BF1.VOXI
 E  BF2.VOXI
 E  BF3.VOXI
 E  BF4.VOXI
 E  BFH.WTXI
 E  BFL.WTXI
 E  BFN.WTXI
 E  BF1.WTXI
 E  BF2.WTXI
 E  BF3.WTXI
 E  BF4.WTXI
 E  CMC...XI
 E  COF.BBXI
 E  CUT.BBXI
 E  COF.BCXI
 E  CUT.BCXI
 E  CUT.BDXI
 E  CUT.BFXI
 E  CUT.BGXI
 E  COF.BHXI
 E  CUT.BHXI
 E  COF.CAXI
 E  CON.CGXI
 E  CUT.CTXI
 E  CON.CUXI
 E  CON.CXXI
 E  CTK.CXXI
 E  CTK.C3XI
 E  CTK.C4XI
 E  CON.DHXI
 E  CUT.DHXI
 E  CUT.DLXI
 E  CTK.DPXI
 E  CUT.DPXI
 E  CUT.DSXI
 E  CON.FDXI
 E  CON.FGXI
 E  CON.FHXI
 E  CON.FNXI
 E  C...FOXI
 E  CTK.FOXI
 E  CON.F1XI
 E  CON.F2XI
 E  CUT.GTXI
 E  CON.GXXI
 E  COF.HOXI
 E  CON.HOXI
 E  CTK.HOXI
 E  CON.HPXI
 E  CTK.HPXI
 E  COF.HSXI
 E  CON.HSXI
 E  CTK.HSXI
 E  CON.HYXI
 E  CUT.KWXI
 E  C...LEXI
 E  CON.LEXI
 E  CTK.LEXI
 E  CON.L1XI
 E  CTK.MMXI
 E  CON.M1XI
 E  CON.M2XI
 E  CON.M3XI
 E  CON.NCXI
 E  CON.NFXI
 E  CON.NNXI
 E  CON.NPXI
 E  CON.NRXI
 E  CON.NSXI
 E  C...OFXI
 E  COF.OFXI
 E  CON.OPXI
 E  CUT.PDXI
 E  CON.PHXI
 E  CTK.PHXI
 E  COF.POXI
 E  CON.POXI
 E  COF.PPXI
 E  CTK.PSXI
 E  C...RAXI
 E  CON.RAXI
 E  CTK.RAXI
 E  C...RDXI
 E  COF.RDXI
 E  CON.RDXI
 E  CUT.RSXI
 E  CUT.RTXI
 E  COF.SFXI
 E  CON.SGXI
 E  COF.SHXI
 E  CON.SHXI
 E  CTK.SHXI
 E  COF.SKXI
 E  CON.SKXI
 E  C...STXI
 E  CON.STXI
 E  CTK.STXI
 E  CON.SWXI
 E  COF.TAXI
 E  CUT.TGXI
 E  CUT.TWXI
 E  COF.WTXI
 E  KFG.PRXI
 E  KLQ.PRXI
 E  UBW...XI
 E  UCW...XI
 E  UDW...XI
 E  UFF...XI
 E  UFS...XI
 E  UHE...XI
 E  UKW...XI
 E  UMC...XI
 E  UOH...XI
 E  UOL...XI
 E  UOM...XI
 E  UPB...XI
 E  UPC...XI
 E  UPD...XI
 E  UPH...XI
 E  UPK...XI
 E  UPL...XI
 E  UPM...XI
 E  UPT...XI
 E  UPW...XI
 E  URW...XI
 E  USH...XI
 E  USL...XI
 E  USM...XI
 E  UTW...XI
 L  XGS.FHXI
 L  XSU.FHXI
 L  XGS.FLXI
 L  XSL.FLXI
 L  XSU.FLXI
 L  XGS.FNXI
 L  XSU.FNXI
 L  XGS.F1XI
 L  XSL.F1XI
 L  XSU.F1XI
 L  XGS.F2XI
 L  XSL.F2XI
 L  XSU.F2XI
 L  XGS.F3XI
 L  XSL.F3XI
 L  XSU.F3XI
 L  XGS.F4XI
 L  XSL.F4XI
 L  XSU.F4XI
 N  FAT0..J.
COLUMNS
    BNP.FHXI  BNP...XI           -1.   BFH...XI         .0929
    BNP.FHXI  BFH.VOXI       1.24425   BFH.WTXI            1.
    BNP.FHXI  C...FOXI       -1.3972
    BDC.FLXI  BDC...XI           -1.   BFL...XI         .0817
    BDC.FLXI  BFL.VOXI       1.08413   BFL.WTXI            1.
    BDC.FLXI  XSU.FLXI           .01   C...FOXI       -1.5995
    BDF.FLXI  BDF...XI           -1.   BFL...XI           -.5
    BDF.FLXI  BFL.VOXI            2.   BFL.WTXI            1.
    BDF.FLXI  XSU.FLXI            .9   XSL.FLXI            .3
    BDF.FLXI  C...FOXI        -1.812
    BDH.FLXI  BDH...XI           -1.   BFL...XI         .0532
    BDH.FLXI  BFL.VOXI       1.06383   BFL.WTXI            1.
    BDH.FLXI  XSU.FLXI           .01   C...FOXI       -1.5995
    BD1.FLXI  BD1...XI           -1.   BFL...XI         .0083
    BD1.FLXI  BFL.VOXI       1.00705   BFL.WTXI            1.
    BD1.FLXI  XSU.FLXI           .22   C...FOXI       -1.5995
    BGF.FLXI  BGF...XI           -1.   BFL...XI        -.1227
    BGF.FLXI  BFL.VOXI         .8547   BFL.WTXI            1.
    BGF.FLXI  XSU.FLXI           1.1   C...FOXI        -1.812
    BG1.FLXI  BG1...XI           -1.   BFL...XI         -.114
    BG1.FLXI  BFL.VOXI        .88028   BFL.WTXI            1.
    BG1.FLXI  XSU.FLXI          1.02   C...FOXI       -2.3923
    BHF.FLXI  BHF...XI           -1.   BFL...XI        -.1079
    BHF.FLXI  BFL.VOXI        .87719   BFL.WTXI            1.
    BHF.FLXI  XSU.FLXI            1.   C...FOXI        -1.812
    BHG.FLXI  BHG...XI           -1.   BFL...XI        -.0484
    BHG.FLXI  BFL.VOXI        .95147   BFL.WTXI            1.
    BHG.FLXI  XSU.FLXI            .2   C...FOXI       -1.5995
    BHH.FLXI  BHH...XI           -1.   BFL...XI           -.5
    BHH.FLXI  BFL.VOXI        .90909   BFL.WTXI            1.
    BHH.FLXI  XSU.FLXI           .01   XSL.FLXI            .4
    BHH.FLXI  C...FOXI       -1.5995
    BHN.FLXI  BHN...XI           -1.   BFL...XI        -.0007
    BHN.FLXI  BFL.VOXI          .998   BFL.WTXI            1.
    BHN.FLXI  XSU.FLXI           .08   C...FOXI       -1.5995
    BHR.FLXI  BHR...XI           -1.   BFL...XI           -.6
    BHR.FLXI  BFL.VOXI          1.25   BFL.WTXI            1.
    BHR.FLXI  XSU.FLXI          .083   XSL.FLXI            2.
    BHR.FLXI  C...FOXI       -1.5995
    BH1.FLXI  BH1...XI           -1.   BFL...XI        -.0057
    BH1.FLXI  BFL.VOXI         .9901   BFL.WTXI            1.
    BH1.FLXI  XSU.FLXI           .32   C...FOXI       -1.5995
    BP1.FLXI  BP1...XI           -1.   BFL...XI        -.1471
    BP1.FLXI  BFL.VOXI        .86505   BFL.WTXI            1.
    BP1.FLXI  XSU.FLXI           1.3   C...FOXI       -2.3923
    BNF.FNXI  C...FOXI       -1.3972
    BNH.FNXI  C...FOXI       -1.3972
    BN1.FNXI  C...FOXI       -1.3972
    BDC.F1XI  BDC...XI           -1.   BF1...XI         .0817
    BDC.F1XI  BF1.VOXI       1.08413   BF1.WTXI            1.
    BDC.F1XI  XSU.F1XI           .01   C...FOXI       -1.5995
    BDF.F1XI  BDF...XI           -1.   BF1...XI           -.5
    BDF.F1XI  BF1.VOXI            2.   BF1.WTXI            1.
    BDF.F1XI  XSU.F1XI            .9   XSL.F1XI            .3
    BDF.F1XI  C...FOXI        -1.812
    BDH.F1XI  BDH...XI           -1.   BF1...XI         .0532
    BDH.F1XI  BF1.VOXI       1.06383   BF1.WTXI            1.
    BDH.F1XI  XSU.F1XI           .01   C...FOXI       -1.5995
    BD1.F1XI  BD1...XI           -1.   BF1...XI         .0083
    BD1.F1XI  BF1.VOXI       1.00705   BF1.WTXI            1.
    BD1.F1XI  XSU.F1XI           .22   C...FOXI       -1.5995
    BGF.F2XI  C...FOXI        -1.812
    BG1.F2XI  C...FOXI       -2.3923
    BHF.F2XI  C...FOXI        -1.812
    BHG.F2XI  BHG...XI           -1.   BF2...XI        -.0484
    BHG.F2XI  BF2.VOXI        .95147   BF2.WTXI            1.
    BHG.F2XI  XSU.F2XI            .2   C...FOXI       -1.5995
    BHH.F2XI  C...FOXI       -1.5995
    BHN.F2XI  BHN...XI           -1.   BF2...XI        -.0007
    BHN.F2XI  BF2.VOXI          .998   BF2.WTXI            1.
    BHN.F2XI  XSU.F2XI           .08   C...FOXI       -1.5995
    BHR.F2XI  BHR...XI           -1.   BF2...XI           -.6
    BHR.F2XI  BF2.VOXI          1.25   BF2.WTXI            1.
    BHR.F2XI  XSU.F2XI          .083   XSL.F2XI            2.
    BHR.F2XI  C...FOXI       -1.5995
    BH1.F2XI  C...FOXI       -1.5995
    BP1.F2XI  C...FOXI       -2.3923
    BGF.F3XI  C...FOXI        -1.812
    BG1.F3XI  C...FOXI       -2.3923
    BHF.F3XI  C...FOXI        -1.812
    BHG.F3XI  BHG...XI           -1.   BF3...XI        -.0484
    BHG.F3XI  BF3.VOXI        .95147   BF3.WTXI            1.
    BHG.F3XI  XSU.F3XI            .2   C...FOXI       -1.5995
    BHH.F3XI  C...FOXI       -1.5995
    BHN.F3XI  BHN...XI           -1.   BF3...XI        -.0007
    BHN.F3XI  BF3.VOXI          .998   BF3.WTXI            1.
    BHN.F3XI  XSU.F3XI           .08   C...FOXI       -1.5995
    BHP.F3XI  BHP...XI           -1.   BF3...XI           -.6
    BHP.F3XI  BF3.VOXI        .83333   BF3.WTXI            1.
    BHP.F3XI  XSU.F3XI          .083   XSL.F3XI            3.
    BHP.F3XI  C...FOXI        -1.812
    BHR.F3XI  BHR...XI           -1.   BF3...XI           -.6
    BHR.F3XI  BF3.VOXI          1.25   BF3.WTXI            1.
    BHR.F3XI  XSU.F3XI          .083   XSL.F3XI            2.
    BHR.F3XI  C...FOXI       -1.5995
    BH1.F3XI  C...FOXI       -1.5995
    BP1.F3XI  C...FOXI       -2.3923
    BSB.F3XI  BSB...XI           -1.   BF3...XI        -.1471
    BSB.F3XI  BF3.VOXI        .83963   BF3.WTXI            1.
    BSB.F3XI  XSU.F3XI           1.3   XSL.F3XI           .73
    BSB.F3XI  C...FOXI        -1.812
    BDC.F4XI  C...FOXI       -1.5995
    BDF.F4XI  C...FOXI        -1.812
    BDH.F4XI  C...FOXI       -1.5995
    BD1.F4XI  C...FOXI       -1.5995
    BGF.F4XI  C...FOXI        -1.812
    BG1.F4XI  C...FOXI       -2.3923
    BHF.F4XI  C...FOXI        -1.812
    BHG.F4XI  BHG...XI           -1.   BF4...XI        -.0484
    BHG.F4XI  BF4.VOXI        .95147   BF4.WTXI            1.
    BHG.F4XI  XSU.F4XI            .2   C...FOXI       -1.5995
    BHH.F4XI  C...FOXI       -1.5995
    BHN.F4XI  BHN...XI           -1.   BF4...XI        -.0007
    BHN.F4XI  BF4.VOXI          .998   BF4.WTXI            1.
    BHN.F4XI  XSU.F4XI           .08   C...FOXI       -1.5995
    BHR.F4XI  BHR...XI           -1.   BF4...XI           -.6
    BHR.F4XI  BF4.VOXI          1.25   BF4.WTXI            1.
    BHR.F4XI  XSU.F4XI          .083   XSL.F4XI            2.
    BHR.F4XI  C...FOXI       -1.5995
    BH1.F4XI  C...FOXI       -1.5995
    BP1.F4XI  C...FOXI       -2.3923
    DFH...XI  BFH...XI           -1.   FAT0..J.           -1.
    DFN...XI  BFN...XI           -1.   FAT0..J.           -1.
    DF1...XI  BF1...XI           -1.   FAT0..J.           -1.
    DF2...XI  BF2...XI           -1.   FAT0..J.           -1.
    DF3...XI  BF3...XI           -1.   FAT0..J.           -1.
    DF4...XI  BF4...XI           -1.   FAT0..J.           -1.
    DL2...XI  BL2...XI           -1.   FAT0..J.           -1.
    DL3...XI  BL3...XI           -1.   FAT0..J.           -1.
    DL4...XI  BL4...XI           -1.   FAT0..J.           -1.
    DMM...XI  BMM...XI           -1.   FAT0..J.           -1.
    DMM...XI  CTK.MMXI           -1.
    DPS...XI  BPS...XI           -1.   FAT0..J.           -1.
    DPS...XI  CTK.PSXI           -1.
    EAR...XI  BAR...XI            1.   TIE.MRAR          90.9
    EMR...XI  BMR...XI            1.   FAT0..J.       35.8586
    EMR...XI  TIE.MRAR          -9.1
    IMC...XI  CMC...XI            1.   FAT0..J.       2.29976
    IOF.BBXI  COF.BBXI            1.   FAT0..J.       3.34313
    IUT.BBXI  CUT.BBXI            1.   FAT0..J.      24.86637
    IOF.BCXI  COF.BCXI            1.   FAT0..J.        .38919
    IUT.BCXI  CUT.BCXI            1.   FAT0..J.       2.64336
    IUT.BDXI  CUT.BDXI            1.
    IUT.BFXI  CUT.BFXI            1.   FAT0..J.        .08031
    IUT.BGXI  CUT.BGXI            1.   FAT0..J.       1.25952
    IOF.BHXI  COF.BHXI            1.   FAT0..J.       4.72575
    IUT.BHXI  CUT.BHXI            1.   FAT0..J.      86.86447
    IOF.CAXI  COF.CAXI            1.   FAT0..J.       3.57535
    ION.CGXI  CON.CGXI            1.   FAT0..J.      10.71777
    IUT.CTXI  CUT.CTXI            1.   FAT0..J.        .01064
    ION.CUXI  CON.CUXI            1.   FAT0..J.      11.32657
    ION.CXXI  CON.CXXI            1.   FAT0..J.       1.78087
    ITK.CXXI  CTK.CXXI            1.   FAT0..J.       1.01103
    ITK.C3XI  CTK.C3XI            1.   FAT0..J.      35.83179
    ITK.C4XI  CTK.C4XI            1.   FAT0..J.       8.10093
    ION.DHXI  CON.DHXI            1.   FAT0..J.       3.69067
    IUT.DHXI  CUT.DHXI            1.   FAT0..J.       2.61058
    IUT.DLXI  CUT.DLXI            1.   FAT0..J.       2.24091
    ITK.DPXI  CTK.DPXI            1.   FAT0..J.      52.42278
    IUT.DPXI  CUT.DPXI            1.   FAT0..J.       3.79887
    IUT.DSXI  CUT.DSXI            1.   FAT0..J.        .24711
    ION.FDXI  CON.FDXI            1.   FAT0..J.       2.05802
    ION.FGXI  CON.FGXI            1.   FAT0..J.       2.28247
    ION.FHXI  CON.FHXI            1.   FAT0..J.       2.75445
    ION.FNXI  CON.FNXI            1.   FAT0..J.       1.54625
    ITK.FOXI  CTK.FOXI            1.   FAT0..J.        .52736
    ION.F1XI  CON.F1XI            1.   FAT0..J.      40.54475
    ION.F2XI  CON.F2XI            1.   FAT0..J.      32.44673
    IUT.GTXI  CUT.GTXI            1.   FAT0..J.      13.81608
    ION.GXXI  CON.GXXI            1.   FAT0..J.     335.35491
    IOF.HOXI  COF.HOXI            1.   FAT0..J.       1.66444
    ION.HOXI  CON.HOXI            1.   FAT0..J.      19.59433
    ITK.HOXI  CTK.HOXI            1.   FAT0..J.        .19963
    ION.HPXI  CON.HPXI            1.   FAT0..J.     146.66831
    ITK.HPXI  CTK.HPXI            1.   FAT0..J.       1.19412
    IOF.HSXI  COF.HSXI            1.   FAT0..J.         .4855
    ION.HSXI  CON.HSXI            1.   FAT0..J.       29.9668
    ITK.HSXI  CTK.HSXI            1.   FAT0..J.        .18299
    ION.HYXI  CON.HYXI            1.   FAT0..J.       8.56792
    IUT.KWXI  CUT.KWXI            1.   FAT0..J.        .00303
    ION.LEXI  CON.LEXI            1.   FAT0..J.        .64005
    ITK.LEXI  CTK.LEXI            1.   FAT0..J.        .68779
    ION.L1XI  CON.L1XI            1.   FAT0..J.       4.25757
    ITK.MMXI  CTK.MMXI            1.   FAT0..J.      36.70052
    ION.M1XI  CON.M1XI            1.   FAT0..J.       7.64337
    ION.M2XI  CON.M2XI            1.   FAT0..J.       6.19878
    ION.M3XI  CON.M3XI            1.   FAT0..J.       4.54615
    ION.NCXI  CON.NCXI            1.   FAT0..J.       10.3509
    ION.NFXI  CON.NFXI            1.   FAT0..J.       1.87473
    ION.NNXI  CON.NNXI            1.   FAT0..J.       3.69496
    ION.NPXI  CON.NPXI            1.   FAT0..J.       3.83929
    ION.NRXI  CON.NRXI            1.   FAT0..J.       8.89789
    ION.NSXI  CON.NSXI            1.   FAT0..J.        2.7243
    IOF.OFXI  COF.OFXI            1.   FAT0..J.       1.25224
    ION.OPXI  CON.OPXI            1.   FAT0..J.       5.52621
    IUT.PDXI  CUT.PDXI            1.   FAT0..J.        .12759
    ION.PHXI  CON.PHXI            1.   FAT0..J.        .41798
    ITK.PHXI  CTK.PHXI            1.   FAT0..J.        .14539
    IOF.POXI  COF.POXI            1.   FAT0..J.        .96767
    ION.POXI  CON.POXI            1.   FAT0..J.       7.02784
    IOF.PPXI  COF.PPXI            1.   FAT0..J.        .77973
    ITK.PSXI  CTK.PSXI            1.   FAT0..J.        .52005
    ION.RAXI  CON.RAXI            1.   FAT0..J.       6.72031
    ITK.RAXI  CTK.RAXI            1.   FAT0..J.       1.73973
    IOF.RDXI  COF.RDXI            1.   FAT0..J.        .00741
    ION.RDXI  CON.RDXI            1.   FAT0..J.        .60438
    IUT.RSXI  CUT.RSXI            1.   FAT0..J.        .12671
    IUT.RTXI  CUT.RTXI            1.   FAT0..J.        .07221
    IOF.SFXI  COF.SFXI            1.   FAT0..J.       1.54999
    ION.SGXI  CON.SGXI            1.   FAT0..J.       2.74321
    IOF.SHXI  COF.SHXI            1.   FAT0..J.        .00907
    ION.SHXI  CON.SHXI            1.   FAT0..J.       1.62408
    ITK.SHXI  CTK.SHXI            1.   FAT0..J.        .27641
    IOF.SKXI  COF.SKXI            1.   FAT0..J.        .20275
    ION.SKXI  CON.SKXI            1.   FAT0..J.      21.07328
    ION.STXI  CON.STXI            1.   FAT0..J.      52.06134
    ITK.STXI  CTK.STXI            1.   FAT0..J.       2.13127
    ION.SWXI  CON.SWXI            1.   FAT0..J.        .21951
    IOF.TAXI  COF.TAXI            1.   FAT0..J.        .37688
    IUT.TGXI  CUT.TGXI            1.   FAT0..J.        .00408
    IUT.TWXI  CUT.TWXI            1.   FAT0..J.        .00031
    IOF.WTXI  COF.WTXI            1.   FAT0..J.        .94443
    KLQ.PRXI  KLQ.PRXI            1.
    MFG.FFXI  BFF...XI            1.   BFG...XI           -1.
    MFI.FFXI  BFF...XI            1.   BFI...XI           -1.
    MBT.FGXI  BBT...XI           -1.   BFG...XI        1.1391
    MBT.FGXI  CUT.DPXI           -1.   CTK.DPXI          -.25
    MBT.FGXI  USL...XI        -.3422
    MFX.FGXI  BFX...XI           -1.   BFG...XI            .9
    MFX.FGXI  BFS...XI            .1   CUT.DLXI           -1.
    MFX.FGXI  CUT.DHXI         -.146
    MHB.FGXI  BHB...XI           -1.   BFG...XI         1.227
    MHB.FGXI  CON.DHXI           -1.   CUT.DHXI           -1.
    MHB.FGXI  UCW...XI        -.4618   UKW...XI        -.5549
    MHB.FGXI  UTW...XI       -73.333
    MPR.FGXI  BPR...XI           -1.   BFG...XI        1.1546
    MPR.FGXI  CUT.DPXI           -1.   CTK.DPXI           -1.
    MPR.FGXI  USL...XI        -.3422
    MIF.FIXI  CUT.DLXI           -.5   CUT.DHXI         -.146
    MIF.FIXI  BIF...XI           -1.   BFI...XI            1.
    MIL.FIXI  CUT.DLXI           -.5   CUT.DHXI         -.146
    MIL.FIXI  BIL...XI           -1.   BFI...XI            1.
    MIP.FIXI  CUT.DLXI           -.5   CUT.DHXI         -.146
    MIP.FIXI  BIP...XI           -1.   BFI...XI            1.
    MIS.FIXI  CUT.DLXI           -.5   CUT.DHXI         -.146
    MIS.FIXI  BIS...XI           -1.   BFI...XI            1.
    MC2.FPXI  CUT.DHXI           -1.   BC2...XI           -1.
    MC2.FPXI  BFP...XI         1.227
    MC3.FPXI  CUT.DHXI           -1.   BC3...XI           -1.
    MC3.FPXI  BFP...XI        1.1546
    MC4.FPXI  CUT.DHXI           -1.   BC4...XI           -1.
    MC4.FPXI  BFP...XI        1.1391
    M6H.FPXI  CUT.DHXI           -1.   B6H...XI           -1.
    M6H.FPXI  BFP...XI        2.9916
    MFG.FSXI  BFS...XI            1.   BFG...XI           -1.
    MFI.FSXI  BFS...XI            1.   BFI...XI           -1.
    MH1.G1XI  BG1...XI            1.   BH1...XI           -1.
    MHB.L2XI  BHB...XI           -1.   BL2...XI         1.279
    MHB.L2XI  UKW...XI        -1.844   UTW...XI      -243.687
    MHB.L2XI  UCW...XI        -1.534   CON.SGXI           -1.
    MPR.L3XI  BPR...XI           -1.   BL3...XI        1.1833
    MPR.L3XI  UKW...XI        -1.309   CTK.C3XI           -1.
    MBT.L4XI  BBT...XI           -1.   BL4...XI         1.164
    MBT.L4XI  UKW...XI        -1.527   CTK.C4XI           -1.
    PAD.BCXI  BAD...XI           -1.   BCE...XI            1.
    PAD.BCXI  COF.BCXI       -26.902
    PAN.BCXI  BAN...XI           -1.   BCE...XI            1.
    PAN.BCXI  COF.BCXI       -31.677
    PAS.BCXI  BAS...XI           -1.   BCE...XI            1.
    PAS.BCXI  COF.BCXI      -27.6428
    PYD.BCXI  USL...XI        -.8393   UPB...XI      -11.5607
    PYD.BCXI  UMC...XI        -1.881   UPW...XI        -.7283
    PYD.BCXI  UPH...XI       11.5607   UKW...XI       -111.45
    PYD.BCXI  BYD...XI           -1.   BCE...XI         .1394
    PYD.BCXI  BWS...XI         .0876   COF.BCXI       -5.9552
    PYD.BCXI  CUT.BCXI       -13.544
    PYN.BCXI  USL...XI        -.8393   UPB...XI      -11.4943
    PYN.BCXI  UMC...XI         -1.97   UPW...XI        -.7628
    PYN.BCXI  UPH...XI       11.4943   UKW...XI       -111.45
    PYN.BCXI  BYN...XI           -1.   BCE...XI          .146
    PYN.BCXI  BWS...XI        .09175   COF.BCXI        -5.921
    PYN.BCXI  CUT.BCXI       -13.466
    PYS.BCXI  USL...XI        -.2127   UPB...XI      -11.5607
    PYS.BCXI  UMC...XI        -1.881   UPW...XI        -.7283
    PYS.BCXI  UPH...XI       11.5607   UKW...XI        -124.2
    PYS.BCXI  BYS...XI           -1.   BCE...XI         .1394
    PYS.BCXI  BWS...XI         .0876   COF.BCXI      -16.8522
    PYS.BCXI  CUT.BCXI       -13.544
    PFS.BGXI  UFS...XI           -1.   UKW...XI      -118.106
    PFS.BGXI  UPB...XI      -12.0337   UPH...XI       12.0337
    PFS.BGXI  USL...XI        -.7858   CUT.BGXI      -13.2371
    PAD.BHXI  BCE...XI            1.   COF.BHXI       -1.0795
    PAD.BHXI  BAD...XI           -1.
    PAN.BHXI  BAN...XI           -1.   BCE...XI            1.
    PAN.BHXI  COF.BHXI       -1.2711
    PAS.BHXI  BCE...XI            1.   COF.BHXI       -1.1619
    PAS.BHXI  BAS...XI           -1.
    PYD.BHXI  UKW...XI       -111.49   UPB...XI        -7.824
    PYD.BHXI  USL...XI        -.8399   UPH...XI         7.824
    PYD.BHXI  UPW...XI        -.6047   UMC...XI        -1.559
    PYD.BHXI  UHE...XI         .2455   BCE...XI        .13553
    PYD.BHXI  CUT.BHXI       -.24552   COF.BHXI       -.24552
    PYD.BHXI  BYD...XI           -1.
    PYN.BHXI  UKW...XI      -110.867   UPB...XI         -7.78
    PYN.BHXI  USL...XI        -.8352   UPH...XI          7.78
    PYN.BHXI  UPW...XI       -.60132   UMC...XI       -1.9716
    PYN.BHXI  UHE...XI         .2441   BYN...XI           -1.
    PYN.BHXI  BCE...XI        .14551   CUT.BHXI        -.2441
    PYN.BHXI  COF.BHXI        -.2441
    PYS.BHXI  UKW...XI      -123.229   UPB...XI        -7.824
    PYS.BHXI  USL...XI        -.2126   UPH...XI         7.824
    PYS.BHXI  UPW...XI        -.6047   UMC...XI        -1.559
    PYS.BHXI  UHE...XI         .2455   BCE...XI        .13553
    PYS.BHXI  CUT.BHXI       -.24552   COF.BHXI        -.6624
    PYS.BHXI  BYS...XI           -1.
    PCC.CGXI  UKW...XI           -2.   UTW...XI        -230.7
    PCC.CGXI  USL...XI         -1.72   UOM...XI         2.169
    PCC.CGXI  UBW...XI        -2.169   UCW...XI         -26.7
    PCC.CGXI  UPW...XI          -.68   UFF...XI        -.0262
    PCC.CGXI  BFX...XI          .661   BPS...XI         .0723
    PCC.CGXI  BWS...XI            .5   BHS...XI         .0029
    PCC.CGXI  BMM...XI         .0046   C...STXI        -.0723
    PCC.CGXI  C...RAXI        -.0046   CON.CGXI        -1.113
    PCC.CGXI  BCC...XI           -1.
    PCF.CGXI  UKW...XI           -2.   UTW...XI        -214.7
    PCF.CGXI  USL...XI        -1.685   UOM...XI         2.095
    PCF.CGXI  UBW...XI        -2.095   UCW...XI         -25.8
    PCF.CGXI  UPW...XI          -.68   UFF...XI        -.0253
    PCF.CGXI  BFX...XI          .672   BPS...XI         .0547
    PCF.CGXI  BWS...XI            .5   BCF...XI           -1.
    PCF.CGXI  BHS...XI         .0029   BMM...XI         .0046
    PCF.CGXI  C...STXI        -.0547   C...RAXI        -.0046
    PCF.CGXI  CON.CGXI       -1.0912
    PCK.CGXI  UKW...XI       -10.525   UTW...XI       -253.15
    PCK.CGXI  USL...XI         -1.32   UOL...XI         1.615
    PCK.CGXI  UBW...XI        -1.615   UCW...XI         -14.9
    PCK.CGXI  UDW...XI          -.33   BCK...XI           -1.
    PCK.CGXI  BFX...XI          .663   BPS...XI         .0325
    PCK.CGXI  BWS...XI           .17   BHS...XI          .002
    PCK.CGXI  BMM...XI         .0015   C...STXI        -.0325
    PCK.CGXI  C...RAXI        -.0015   CON.CGXI           -1.
    PAN.COXI  UKW...XI        -12.91   USM...XI        -.1975
    PAN.COXI  USL...XI        -.1544   UOL...XI         .0996
    PAN.COXI  UBW...XI        -.0996   UCW...XI       -1.3053
    PAN.COXI  BAN...XI           -1.   BCK...XI           -.5
    PAN.COXI  BWS...XI         .1289   BCE...XI           1.5
    PAN.COXI  CON.CXXI           -1.   CTK.CXXI           -1.
    PAN.COXI  COF.CAXI          -1.5
    PYN.COXI  UKW...XI        -12.91   USM...XI        -.1975
    PYN.COXI  USL...XI        -.1544   UOL...XI         .0996
    PYN.COXI  UBW...XI        -.0996   UCW...XI       -1.3053
    PYN.COXI  BYN...XI           -1.   BC2...XI         .0366
    PYN.COXI  BC3...XI         .0038   BC4...XI         .0015
    PYN.COXI  BNF...XI         .0057   BDF...XI         .0223
    PYN.COXI  BHF...XI         .0394   BGF...XI         .0735
    PYN.COXI  BSB...XI         .1533   BCK...XI         .6529
    PYN.COXI  BHS...XI         .0023   BMM...XI         .0011
    PYN.COXI  BWS...XI         .1289   B6H...XI         .0045
    PYN.COXI  C...LEXI         -.011   C...RAXI        -.0011
    PYN.COXI  CON.CXXI           -1.   CTK.CXXI           -1.
    PYN.COXI  C...RDXI        -.0519
    PAN.CRXI  UKW...XI        -12.91   USM...XI         -.223
    PAN.CRXI  USL...XI        -.1544   UOL...XI         .0996
    PAN.CRXI  UBW...XI        -.0996   UCW...XI       -1.4738
    PAN.CRXI  BAN...XI           -1.   BCK...XI           -.5
    PAN.CRXI  BWS...XI         .1917   BCE...XI           1.5
    PAN.CRXI  CON.CXXI        -1.037   CTK.CXXI           -1.
    PAN.CRXI  COF.CAXI          -1.5
    PYN.CRXI  UKW...XI        -12.91   USM...XI         -.223
    PYN.CRXI  USL...XI        -.1544   UOL...XI         .0996
    PYN.CRXI  UBW...XI        -.0996   UCW...XI       -1.4738
    PYN.CRXI  BYN...XI           -1.   BC2...XI         .0437
    PYN.CRXI  BC3...XI         .0045   BC4...XI         .0015
    PYN.CRXI  BNF...XI         .0063   BDF...XI          .025
    PYN.CRXI  BHF...XI         .0441   BGF...XI         .0826
    PYN.CRXI  BCK...XI         .7861   BHS...XI         .0023
    PYN.CRXI  BMM...XI         .0011   BWS...XI         .1917
    PYN.CRXI  B6H...XI         .0055   C...LEXI        -.0123
    PYN.CRXI  C...RAXI        -.0011   C...RDXI        -.0519
    PYN.CRXI  CON.CXXI        -1.037   CTK.CXXI           -1.
    PHS.CUXI  UKW...XI       -88.454   UBW...XI         -3.17
    PHS.CUXI  UFF...XI         -.061   UOL...XI          3.17
    PHS.CUXI  BHS...XI        -1.064   BPS...XI            1.
    PHS.CUXI  C...STXI          -.05   CON.CUXI           -1.
    PAB.FGXI  UBW...XI        -.0944   UCW...XI       -1.1371
    PAB.FGXI  UFF...XI        -.0025   UOL...XI         .0944
    PAB.FGXI  USL...XI        -.0126   UKW...XI       -5.9631
    PAB.FGXI  BAB...XI           -1.   BG1...XI        -.0245
    PAB.FGXI  BWS...XI         .0076   BYG...XI         .0245
    PAB.FGXI  BAG...XI            1.   CON.FGXI           -1.
    PPX.FGXI  UBW...XI        -.0944   UCW...XI       -1.1371
    PPX.FGXI  UFF...XI        -.0025   UOL...XI         .0944
    PPX.FGXI  USL...XI        -.0126   UKW...XI       -5.9631
    PPX.FGXI  BG1...XI        -.0245   BPX...XI           -1.
    PPX.FGXI  BWS...XI         .0076   BYG...XI        1.0245
    PPX.FGXI  CON.FGXI           -1.   KFG.PRXI           -1.
    P...FOXI  USM...XI         -.003   USL...XI        -.0136
    P...FOXI  UKW...XI        -.0006   C...FOXI            1.
    P...FOXI  CTK.FOXI           -1.
    P...F1XI  UKW...XI      -27.8678   UFF...XI       -1.1224
    P...F1XI  UHE...XI            1.   CON.F1XI           -1.
    PAD.GXXI  BAD...XI           -1.   CON.GXXI           -1.
    PAG.GXXI  CON.GXXI           -1.   BAG...XI           -1.
    PAH.GXXI  CON.GXXI           -1.   BAH...XI           -1.
    PAN.GXXI  BAN...XI           -1.   CON.GXXI           -1.
    PYD.GXXI  BYD...XI           -1.   CON.GXXI           -1.
    PYG.GXXI  CON.GXXI           -1.   BYG...XI           -1.
    PYH.GXXI  CON.GXXI           -1.   BYH...XI           -1.
    PYN.GXXI  BYN...XI           -1.   CON.GXXI           -1.
    PAD.HOXI  BAD...XI           -1.   BCE...XI            1.
    PAD.HOXI  CON.HOXI           -1.   COF.HOXI           -1.
    PAD.HOXI  CTK.HOXI           -1.
    PAN.HOXI  BCE...XI            1.   CON.HOXI      -1.02188
    PAN.HOXI  COF.HOXI       -1.0162   CTK.HOXI       -1.0094
    PAN.HOXI  BAN...XI           -1.
    PYD.HOXI  UKW...XI         -152.   UTW...XI        -247.1
    PYD.HOXI  UCW...XI        -20.92   UDW...XI        -2.184
    PYD.HOXI  USL...XI         -1.23   UMC...XI         -1.15
    PYD.HOXI  BYD...XI           -1.   BG1...XI        -.0251
    PYD.HOXI  BO2...XI       -1.0127   B9H...XI         .2082
    PYD.HOXI  BHB...XI          .031   BPS...XI         .0217
    PYD.HOXI  BWS...XI         .5268   BCE...XI           .01
    PYD.HOXI  CON.HOXI           -1.   COF.HOXI           -1.
    PYD.HOXI  CTK.HOXI           -1.
    PYN.HOXI  UKW...XI         -152.   UTW...XI        -247.1
    PYN.HOXI  UCW...XI        -20.92   UDW...XI        -2.184
    PYN.HOXI  USL...XI         -1.23   UMC...XI         -1.15
    PYN.HOXI  BG1...XI        -.0251   BO2...XI       -1.0127
    PYN.HOXI  B9H...XI         .2097   BHB...XI          .031
    PYN.HOXI  BPS...XI         .0202   BWS...XI         .5268
    PYN.HOXI  BCE...XI           .01   CON.HOXI      -1.02188
    PYN.HOXI  COF.HOXI       -1.0162   CTK.HOXI       -1.0094
    PYN.HOXI  BYN...XI           -1.
    PAC.HSXI  BAC...XI           -1.   BCE...XI            1.
    PAC.HSXI  CON.HSXI           -1.   COF.HSXI           -1.
    PAC.HSXI  CTK.HSXI           -1.
    PAF.HSXI  BCE...XI            1.   CON.HSXI           -.9
    PAF.HSXI  COF.HSXI         -1.81   CTK.HSXI           -1.
    PAF.HSXI  BAF...XI           -1.
    PAY.HSXI  BCE...XI            1.   CON.HSXI           -.9
    PAY.HSXI  COF.HSXI         -2.14   CTK.HSXI           -1.
    PAY.HSXI  BAY...XI           -1.
    PIC.HSXI  UKW...XI        -176.9   UTW...XI        -231.1
    PIC.HSXI  UCW...XI        -15.85   UDW...XI        -1.733
    PIC.HSXI  USL...XI        -1.053   UMC...XI         -1.15
    PIC.HSXI  BIC...XI           -1.   BO2...XI        -.7955
    PIC.HSXI  B9H...XI         .1715   BHB...XI         .0376
    PIC.HSXI  BPS...XI         .0487   BWS...XI         .5268
    PIC.HSXI  BCE...XI          .026   CON.HSXI           -1.
    PIC.HSXI  COF.HSXI           -1.   CTK.HSXI           -1.
    PYF.HSXI  UKW...XI        -176.9   UTW...XI        -231.1
    PYF.HSXI  UCW...XI        -15.85   UDW...XI        -1.733
    PYF.HSXI  USL...XI        -1.053   UMC...XI         -1.15
    PYF.HSXI  BO2...XI           -1.   B9H...XI            .2
    PYF.HSXI  BHB...XI         .0376   BPS...XI         .0316
    PYF.HSXI  BWS...XI         .5268   BCE...XI          .026
    PYF.HSXI  CON.HSXI           -.9   COF.HSXI         -1.81
    PYF.HSXI  CTK.HSXI           -1.   BYF...XI           -1.
    PYY.HSXI  UKW...XI        -176.9   UTW...XI        -231.1
    PYY.HSXI  UCW...XI        -15.85   UDW...XI        -1.733
    PYY.HSXI  USL...XI        -1.053   UMC...XI         -1.15
    PYY.HSXI  BO2...XI           -1.   B9H...XI            .2
    PYY.HSXI  BHB...XI         .0376   BPS...XI         .0403
    PYY.HSXI  BWS...XI         .5268   BCE...XI          .026
    PYY.HSXI  CON.HSXI           -.9   COF.HSXI         -2.14
    PYY.HSXI  CTK.HSXI           -1.   BYY...XI           -1.
    PC2.HYXI  UKW...XI      -71.7838   UTW...XI       -130.53
    PC2.HYXI  USM...XI        -.1065   UCW...XI        -5.867
    PC2.HYXI  UMC...XI       -1.1676   BC2...XI           -1.
    PC2.HYXI  BN1...XI        -.0158   BHB...XI        1.0126
    PC2.HYXI  BPR...XI         .0015   CON.HYXI           -1.
    PC3.HYXI  UKW...XI     -225.0204   UTW...XI       -130.53
    PC3.HYXI  USM...XI        -.1065   UCW...XI        -5.867
    PC3.HYXI  UMC...XI       -1.1676   BC3...XI           -1.
    PC3.HYXI  BHB...XI         .1805   BPR...XI         .7863
    PC3.HYXI  BBT...XI         .0222   CON.HYXI           -1.
    PC4.HYXI  UKW...XI     -225.0204   UTW...XI       -130.53
    PC4.HYXI  USM...XI        -.1065   UCW...XI        -5.867
    PC4.HYXI  UMC...XI       -1.1676   BC4...XI           -1.
    PC4.HYXI  BHB...XI         .1003   BPR...XI         .0335
    PC4.HYXI  BBT...XI         .8223   BNH...XI         .0257
    PC4.HYXI  CON.HYXI           -1.
    P6H.HYXI  UKW...XI      -76.0552   UTW...XI     -1426.904
    P6H.HYXI  USM...XI        -.1065   UCW...XI       -62.674
    P6H.HYXI  UMC...XI       -1.1676   B6H...XI           -1.
    P6H.HYXI  B9H...XI         .8703   BHB...XI         .0689
    P6H.HYXI  BWS...XI         .1828   CON.HYXI           -1.
    P...LEXI  UKW...XI         -.724   USM...XI        -.2575
    P...LEXI  USL...XI        -.0018   UCW...XI       -8.6439
    P...LEXI  UFF...XI        -.0001   UMC...XI        -.0054
    P...LEXI  C...LEXI            1.   CON.LEXI           -1.
    P...LEXI  CTK.LEXI           -1.
    PIC.M3XI  UKW...XI        -39.37   UCW...XI         -2.15
    PIC.M3XI  USL...XI         -.091   UHE...XI        -.0541
    PIC.M3XI  BC2...XI        .07085   BC3...XI        .03069
    PIC.M3XI  BC4...XI         .0176   BD1...XI        .21287
    PIC.M3XI  BG1...XI         .0418   BHS...XI        .03223
    PIC.M3XI  BH1...XI        .05688   BIC...XI           -1.
    PIC.M3XI  BMM...XI        .00572   BN1...XI        .16403
    PIC.M3XI  BPX...XI        .30363   BRS...XI          -1.2
    PIC.M3XI  BSS...XI           1.2   BWP...XI         .1647
    PIC.M3XI  BWS...XI          .132   B6H...XI        .01331
    PIC.M3XI  B9H...XI       -.05138   C...RAXI        -.0057
    PIC.M3XI  C...RDXI         -.244   C...LEXI        -.2123
    PIC.M3XI  CON.M3XI           -1.   C...OFXI           -1.
    PIC.M3XI  KLQ.PRXI           -1.   KFG.PRXI        .30363
    PGF.NCXI  UCW...XI         -2.22   UKW...XI         -29.4
    PGF.NCXI  UTW...XI         -53.4   UMC...XI         -3.24
    PGF.NCXI  UFF...XI        -.0503   BC2...XI         .0129
    PGF.NCXI  BC3...XI          .015   BC4...XI         .0352
    PGF.NCXI  BHS...XI         .0088   BWS...XI         .0422
    PGF.NCXI  B9H...XI         -.079   BMM...XI         .0134
    PGF.NCXI  BNH...XI         .2803   BGF...XI           -1.
    PGF.NCXI  CON.NCXI           -1.   C...LEXI        -.0502
    PGF.NCXI  C...RAXI        -.0134   C...RDXI         -.072
    PGF.NCXI  BDC...XI         .6532
    PG1.NCXI  UCW...XI         -2.22   UKW...XI         -29.4
    PG1.NCXI  UTW...XI         -53.4   UMC...XI         -3.24
    PG1.NCXI  UFF...XI        -.0503   BC2...XI         .0129
    PG1.NCXI  BC3...XI          .015   BC4...XI         .0352
    PG1.NCXI  BHS...XI         .0088   BWS...XI         .0422
    PG1.NCXI  B9H...XI        -.0784   BMM...XI         .0134
    PG1.NCXI  BNH...XI         .2803   BG1...XI           -1.
    PG1.NCXI  CON.NCXI           -1.   C...LEXI        -.0502
    PG1.NCXI  C...RAXI        -.0134   C...RDXI         -.072
    PG1.NCXI  BDC...XI         .6532
    PHF.NCXI  UCW...XI         -2.22   UKW...XI         -29.4
    PHF.NCXI  UTW...XI         -53.4   UMC...XI         -3.24
    PHF.NCXI  UFF...XI        -.0503   BC2...XI         .0129
    PHF.NCXI  BC3...XI          .015   BC4...XI         .0352
    PHF.NCXI  BHS...XI         .0088   BWS...XI         .0422
    PHF.NCXI  B9H...XI         -.074   BMM...XI         .0134
    PHF.NCXI  BNH...XI         .2803   BHF...XI           -1.
    PHF.NCXI  CON.NCXI           -1.   C...LEXI        -.0502
    PHF.NCXI  C...RAXI        -.0134   C...RDXI         -.072
    PHF.NCXI  BDC...XI         .6532
    PHH.NCXI  UCW...XI         -2.22   UKW...XI         -29.4
    PHH.NCXI  UTW...XI         -53.4   UMC...XI         -3.24
    PHH.NCXI  UFF...XI        -.0503   BC2...XI         .0129
    PHH.NCXI  BC3...XI          .015   BC4...XI         .0352
    PHH.NCXI  BHS...XI         .0088   BWS...XI         .0422
    PHH.NCXI  B9H...XI         -.042   BMM...XI         .0134
    PHH.NCXI  BNH...XI         .2803   BHH...XI           -1.
    PHH.NCXI  CON.NCXI           -1.   C...LEXI        -.0502
    PHH.NCXI  C...RAXI        -.0134   C...RDXI         -.072
    PHH.NCXI  BDC...XI         .6532
    PH1.NCXI  UCW...XI         -2.22   UKW...XI         -29.4
    PH1.NCXI  UTW...XI         -53.4   UMC...XI         -3.24
    PH1.NCXI  UFF...XI        -.0503   BC2...XI         .0129
    PH1.NCXI  BC3...XI          .015   BC4...XI         .0352
    PH1.NCXI  BHS...XI         .0088   BWS...XI         .0422
    PH1.NCXI  B9H...XI        -.0521   BMM...XI         .0134
    PH1.NCXI  BNH...XI         .2803   BH1...XI           -1.
    PH1.NCXI  CON.NCXI           -1.   C...LEXI        -.0502
    PH1.NCXI  C...RAXI        -.0134   C...RDXI         -.072
    PH1.NCXI  BDC...XI         .6532
    PP1.NCXI  UCW...XI         -2.22   UKW...XI         -29.4
    PP1.NCXI  UTW...XI         -53.4   UMC...XI         -3.24
    PP1.NCXI  UFF...XI        -.0503   BC2...XI         .0129
    PP1.NCXI  BC3...XI          .015   BC4...XI         .0352
    PP1.NCXI  BHS...XI         .0088   BWS...XI         .0422
    PP1.NCXI  B9H...XI         -.087   BMM...XI         .0134
    PP1.NCXI  BNH...XI         .2803   BP1...XI           -1.
    PP1.NCXI  CON.NCXI           -1.   C...LEXI        -.0502
    PP1.NCXI  C...RAXI        -.0134   C...RDXI         -.072
    PP1.NCXI  BDC...XI         .6532
    PNF.NFXI  UCW...XI        -1.256   UKW...XI        -18.02
    PNF.NFXI  UMC...XI        -1.539   UFF...XI        -.0026
    PNF.NFXI  BMM...XI         .0007   BC2...XI         .0023
    PNF.NFXI  BC3...XI         .0023   BHS...XI        .00459
    PNF.NFXI  BNF...XI           -1.   BNH...XI          .677
    PNF.NFXI  B9H...XI       -.01096   BNP...XI          .323
    PNF.NFXI  CON.NFXI           -1.   C...LEXI         .9977
    PNF.NFXI  C...RAXI        -.0007   C...RDXI        -.0375
    PN1.NFXI  UCW...XI        -1.256   UKW...XI        -18.02
    PN1.NFXI  UMC...XI        -1.539   UFF...XI        -.0026
    PN1.NFXI  BMM...XI         .0007   BC2...XI         .0023
    PN1.NFXI  BC3...XI         .0023   BHS...XI        .00459
    PN1.NFXI  BNH...XI          .677   BN1...XI           -1.
    PN1.NFXI  B9H...XI       -.01096   BNP...XI          .323
    PN1.NFXI  CON.NFXI           -1.   C...LEXI         .9977
    PN1.NFXI  C...RAXI        -.0007   C...RDXI        -.0375
    PNH.NPXI  UCW...XI         -.907   UKW...XI        -29.02
    PNH.NPXI  UMC...XI         -.459   UFF...XI        -.0407
    PNH.NPXI  BC3...XI          .015   BC4...XI          .023
    PNH.NPXI  BHB...XI           .02   BNH...XI           -1.
    PNH.NPXI  B9H...XI          .034   BNP...XI         .9078
    PNH.NPXI  CON.NPXI           -1.   C...LEXI         -.038
    P...OFXI  UKW...XI       -5.6433   UCW...XI        -.3123
    P...OFXI  USL...XI        -.0167   UMC...XI         -.002
    P...OFXI  C...OFXI            1.   COF.OFXI           -1.
    P...OPXI  UTW...XI      -428.564   UCW...XI       -24.459
    P...OPXI  BO2...XI            1.   CON.OPXI           -1.
    PWP.PHXI  UKW...XI       -1.6753   USL...XI        -.0342
    PWP.PHXI  UCW...XI       -1.5815   UMC...XI        -.0725
    PWP.PHXI  BWP...XI           -1.   BWW...XI            1.
    PWP.PHXI  CON.PHXI           -1.   CTK.PHXI           -1.
    PAD.POXI  BAD...XI           -1.   BCE...XI            1.
    PAD.POXI  CON.POXI           -1.   COF.POXI           -1.
    PAG.POXI  BCE...XI            1.   CON.POXI       -1.0806
    PAG.POXI  COF.POXI       -1.1032   BAG...XI           -1.
    PAH.POXI  BAH...XI           -1.   BCE...XI            1.
    PAH.POXI  CON.POXI      -1.00594   COF.POXI       -1.2542
    PAN.POXI  BCE...XI            1.   CON.POXI      -1.02315
    PAN.POXI  COF.POXI      -1.01605   BAN...XI           -1.
    PG1.POXI  UKW...XI         -69.7   UTW...XI         -69.5
    PG1.POXI  UCW...XI          -3.5   USL...XI         -.079
    PG1.POXI  UBW...XI       -1.8523   UOH...XI        1.4972
    PG1.POXI  BG1...XI           -1.   BO2...XI         -.989
    PG1.POXI  BIP...XI         .7726   BPS...XI         .0051
    PG1.POXI  BWS...XI         .0325   BCE...XI           .01
    PG1.POXI  CON.POXI      -1.08765   COF.POXI        -.2514
    PH1.POXI  UKW...XI          -72.   UTW...XI          -70.
    PH1.POXI  UCW...XI          -3.5   USL...XI         -.079
    PH1.POXI  UBW...XI       -1.8523   UOH...XI        1.4972
    PH1.POXI  BO2...XI       -1.0216   BIP...XI         .8173
    PH1.POXI  BPS...XI         .0035   BWS...XI         .0325
    PH1.POXI  BH1...XI           -1.   BCE...XI           .01
    PH1.POXI  CON.POXI       -1.0871   COF.POXI        -.2514
    PYD.POXI  UKW...XI        -64.78   UTW...XI        -62.44
    PYD.POXI  UCW...XI       -3.1946   USL...XI        -.0789
    PYD.POXI  UBW...XI       -1.8523   UOH...XI        1.4972
    PYD.POXI  BYD...XI           -1.   BG1...XI        -.0251
    PYD.POXI  BO2...XI       -1.0127   BIP...XI         .7338
    PYD.POXI  BPS...XI         .0217   BWS...XI         .0325
    PYD.POXI  BCE...XI           .01   CON.POXI           -1.
    PYD.POXI  COF.POXI           -1.
    PYG.POXI  UKW...XI        -70.85   UTW...XI        -62.51
    PYG.POXI  UCW...XI        -3.198   USL...XI        -.0939
    PYG.POXI  UBW...XI       -1.8484   UOH...XI        1.4708
    PYG.POXI  BG1...XI        -.0261   BO2...XI       -1.0148
    PYG.POXI  BIP...XI         .7175   BPS...XI         .0303
    PYG.POXI  BWS...XI         .0138   BCE...XI           .01
    PYG.POXI  CON.POXI       -1.0806   COF.POXI       -1.1032
    PYG.POXI  BYG...XI           -1.
    PYH.POXI  UKW...XI        -81.14   UTW...XI        -62.63
    PYH.POXI  UCW...XI        -3.205   USL...XI        -.1192
    PYH.POXI  UBW...XI       -1.8617   UOH...XI        1.4738
    PYH.POXI  BG1...XI        -.0282   BO2...XI       -1.0169
    PYH.POXI  BIP...XI          .719   BPS...XI         .0385
    PYH.POXI  BWS...XI         .0177   BYH...XI           -1.
    PYH.POXI  BCE...XI           .01   CON.POXI      -1.00594
    PYH.POXI  COF.POXI       -1.2542
    PYN.POXI  UKW...XI        -64.78   UTW...XI        -62.44
    PYN.POXI  UCW...XI       -3.1946   USL...XI        -.0789
    PYN.POXI  UBW...XI       -1.8523   UOH...XI        1.4972
    PYN.POXI  BG1...XI        -.0251   BO2...XI       -1.0127
    PYN.POXI  BIP...XI         .7389   BPS...XI         .0197
    PYN.POXI  BWS...XI         .0325   BCE...XI           .01
    PYN.POXI  CON.POXI      -1.02315   COF.POXI      -1.01605
    PYN.POXI  BYN...XI           -1.
    PAR.PPXI  UKW...XI       -6.4569   BAR...XI           -1.
    PAR.PPXI  BAC...XI            1.   COF.PPXI           -1.
    PMR.PPXI  UKW...XI       -6.4569   BMR...XI           -1.
    PMR.PPXI  BIC...XI            1.   COF.PPXI           -1.
    P...RAXI  UKW...XI      -258.877   UCW...XI      -79.2522
    P...RAXI  UMC...XI       -5.6863   C...RAXI            1.
    P...RAXI  CON.RAXI           -1.   CTK.RAXI           -1.
    P...RDXI  USL...XI        -.4221   UKW...XI        -9.183
    P...RDXI  UCW...XI       -2.4961   UDW...XI        -.0089
    P...RDXI  UMC...XI        -.1546   C...RDXI            1.
    P...RDXI  CON.RDXI           -1.   COF.RDXI           -1.
    PAD.SFXI  UKW...XI         -144.   UCW...XI           -8.
    PAD.SFXI  UPW...XI           -.1   BAD...XI           -1.
    PAD.SFXI  BAS...XI            1.   BWS...XI            .1
    PAD.SFXI  COF.SFXI           -1.
    PAG.SFXI  UKW...XI         -144.   UCW...XI           -8.
    PAG.SFXI  UPW...XI           -.1   BWS...XI            .1
    PAG.SFXI  COF.SFXI           -1.   BAG...XI           -1.
    PAG.SFXI  BAF...XI            1.
    PYD.SFXI  UKW...XI         -144.   UCW...XI           -8.
    PYD.SFXI  UPW...XI           -.1   BYD...XI           -1.
    PYD.SFXI  BYS...XI            1.   BWS...XI            .1
    PYD.SFXI  COF.SFXI           -1.
    PYG.SFXI  UKW...XI         -144.   UCW...XI           -8.
    PYG.SFXI  UPW...XI           -.1   BWS...XI            .1
    PYG.SFXI  COF.SFXI           -1.   BYG...XI           -1.
    PYG.SFXI  BYF...XI            1.
    PD1.SHXI  UKW...XI       -1.0695   UCW...XI        -.8357
    PD1.SHXI  UFF...XI        -.0022   USL...XI        -.0703
    PD1.SHXI  UTW...XI       -8.8803   UDW...XI        -.0084
    PD1.SHXI  UMC...XI        -.6196   BD1...XI           -1.
    PD1.SHXI  B9H...XI        -.0103   BDH...XI         .9811
    PD1.SHXI  BC2...XI         .0013   BC3...XI         .0001
    PD1.SHXI  BNH...XI         .0124   BHS...XI         .0023
    PD1.SHXI  BMM...XI         .0007   BWS...XI         .0184
    PD1.SHXI  B6H...XI         .0016   CON.SHXI           -1.
    PD1.SHXI  COF.SHXI           -1.   CTK.SHXI           -1.
    PD1.SHXI  C...RDXI        -.0144   C...RAXI        -.0007
    PH1.SHXI  UKW...XI       -1.0695   UCW...XI        -.8357
    PH1.SHXI  UFF...XI        -.0022   USL...XI        -.0703
    PH1.SHXI  UTW...XI       -8.8803   UDW...XI        -.0084
    PH1.SHXI  UMC...XI        -.6196   BH1...XI           -1.
    PH1.SHXI  B9H...XI        -.0103   BHH...XI         .9811
    PH1.SHXI  BC2...XI         .0013   BC3...XI         .0001
    PH1.SHXI  BNH...XI         .0124   BHS...XI         .0023
    PH1.SHXI  BMM...XI         .0007   BWS...XI         .0184
    PH1.SHXI  B6H...XI         .0016   CON.SHXI           -1.
    PH1.SHXI  COF.SHXI           -1.   CTK.SHXI           -1.
    PH1.SHXI  C...RDXI        -.0144   C...RAXI        -.0007
    PSS.SHXI  UKW...XI       -1.0695   UCW...XI        -.8357
    PSS.SHXI  UFF...XI        -.0022   UTW...XI       -8.8803
    PSS.SHXI  UDW...XI        -.0084   UMC...XI        -.6196
    PSS.SHXI  BSS...XI           -1.   BD1...XI        -.0154
    PSS.SHXI  BH1...XI        -.0039   B9H...XI        -.0103
    PSS.SHXI  BRS...XI            1.   BC2...XI         .0013
    PSS.SHXI  BC3...XI         .0001   BNH...XI         .0124
    PSS.SHXI  BHS...XI         .0023   BMM...XI         .0007
    PSS.SHXI  BWS...XI         .0184   B6H...XI         .0016
    PSS.SHXI  CON.SHXI           -1.   COF.SHXI           -1.
    PSS.SHXI  CTK.SHXI           -1.   C...RDXI        -.0144
    PSS.SHXI  C...RAXI        -.0007
    P...STXI  UKW...XI     -1137.988   USL...XI       -2.9706
    P...STXI  UDW...XI      -12.4507   UMC...XI      -41.1737
    P...STXI  C...STXI            1.   CON.STXI           -1.
    P...STXI  CTK.STXI           -1.
    PWS.SWXI  UKW...XI        -.7547   USL...XI        -.3277
    PWS.SWXI  BWS...XI           -1.   BWP...XI            1.
    PWS.SWXI  CON.SWXI           -1.
    PCE.TAXI  UKW...XI        -.5422   UMC...XI        -7.381
    PCE.TAXI  BCE...XI           -1.   COF.TAXI          -1.2
    PWW.WTXI  UKW...XI        -7.955   UMC...XI        -.6206
    PWW.WTXI  UPW...XI         .5362   BWW...XI           -1.
    PWW.WTXI  COF.WTXI          -1.3   CUT.RTXI        -.5878
    PWW.WTXI  CUT.RSXI        -.5878
    UPH.BYXI  UPH...XI            1.
    UPK.BYXI  UPK...XI            1.
    UPL.BYXI  UPL...XI            1.
    UPM.BYXI  UPM...XI            1.
    URW.BYXI  URW...XI            1.
    UPL.GMXI  UPL...XI            1.   UPH...XI           -1.
    UPL.GMXI  CUT.DSXI          -1.1
    UFF.GNXI  UFF...XI            1.   BFF...XI           -1.
    UFS.GNXI  UFS...XI            1.   BFS...XI           -1.
    UMC.GNXI  UMC...XI            1.   CMC...XI           -1.
    UPB.GNXI  UPB...XI            1.   UKW...XI       -1.7328
    UPB.GNXI  UPD...XI        -.5971   CUT.BFXI          -1.1
    UPB.GNXI  USL...XI         -.046   UMC...XI        -.0142
    UPC.GNXI  UPC...XI            1.   UKW...XI        -.3213
    UPC.GNXI  UPW...XI        -.0449   CUT.CTXI          -1.1
    UPC.GNXI  USL...XI        -.0031   UMC...XI        -.0112
    UPD.GNXI  UPD...XI            1.   URW...XI         .1358
    UPD.GNXI  UKW...XI        -.3984   UPW...XI       -1.1358
    UPD.GNXI  CUT.PDXI          -1.1   UMC...XI         -.005
    UPL.GNXI  UPL...XI            1.   UPM...XI           -1.
    UPL.GNXI  CUT.DSXI          -1.1
    UPM.GNXI  UPM...XI            1.   UPH...XI           -1.
    UPM.GNXI  CUT.DSXI          -1.1
    UPT.GNXI  UPT...XI            1.   UKW...XI        -.9091
    UPW.GNXI  UPW...XI            1.   URW...XI           -1.
    UPW.GNXI  UKW...XI        -.1989   CUT.RTXI          -1.1
    UPW.GNXI  UMC...XI        -.0157
    URW.GNXI  URW...XI            1.   UKW...XI       -1.5952
    URW.GNXI  CUT.RSXI          -1.1
    UPH.GOXI  UPH...XI            1.   UOH...XI           -1.
    UPL.GOXI  UPL...XI            1.   UOL...XI           -1.
    UPM.GOXI  UPM...XI            1.   UOM...XI           -1.
    UBW.GPXI  UBW...XI            1.   UPB...XI         -1.15
    UCW.GPXI  UCW...XI            1.   UPC...XI         -1.15
    UDW.GPXI  UDW...XI            1.   UPD...XI          -1.3
    UKW.GPXI  UKW...XI            1.   CUT.KWXI        -1.265
    UKW.GPXI  UPK...XI         -1.15
    USH.GPXI  USH...XI            1.   UPH...XI         -1.15
    USH.GPXI  CUT.DSXI        -1.265
    USL.GPXI  USL...XI            1.   UPL...XI         -1.15
    USL.GPXI  CUT.DSXI        -1.265
    USM.GPXI  USM...XI            1.   UPM...XI         -1.15
    USM.GPXI  CUT.DSXI        -1.265
    UTW.GPXI  UTW...XI            1.   UPT...XI          -1.1
    UFF.GXXI  UFF...XI           -1.   BFF...XI           -1.
    UFF.GXXI  CON.GXXI           -1.
    UFS.GXXI  UFS...XI           -1.   BFS...XI           -1.
    UFS.GXXI  CON.GXXI           -1.
    UPH.GXXI  UPH...XI           -1.   CON.GXXI           -1.
    UPL.GXXI  UPL...XI           -1.   CON.GXXI           -1.
    UPM.GXXI  UPM...XI           -1.   CON.GXXI           -1.
    UPK.G1XI  UPK...XI            1.   UPH...XI        -.0235
    UPK.G1XI  UPM...XI         .0235   CUT.TGXI          -1.1
    UPT.G1XI  UPT...XI            1.   CUT.TWXI          -1.1
    UPT.G1XI  UPH...XI        -.0352   UPM...XI         .0352
    UPT.G1XI  CUT.DSXI        -.0387
    UPK.G2XI  UPK...XI            1.   UPH...XI        -.0082
    UPK.G2XI  UPL...XI         .0082   CUT.TGXI          -1.1
    UPT.G2XI  UPT...XI            1.   CUT.TWXI          -1.1
    UPT.G2XI  UPH...XI        -.0118   UPL...XI         .0118
    UPT.G2XI  CUT.DSXI         -.013
    UPK.G3XI  UPK...XI            1.   UPH...XI        -.0044
    UPK.G3XI  UCW...XI        -.1271   UPB...XI         .0044
    UPK.G3XI  CUT.TGXI          -1.1
    UPT.G3XI  UPT...XI            1.   CUT.TWXI          -1.1
    UPT.G3XI  UPH...XI        -.0057   UCW...XI        -.0883
    UPT.G3XI  CUT.DSXI        -.0063
    UPT.G4XI  UPT...XI            1.   CUT.TWXI          -1.1
    UPT.G4XI  UPM...XI        -.0183   UPL...XI         .0183
    UPT.G4XI  CUT.DSXI        -.0201
    UPT.G5XI  UPT...XI            1.   CUT.TWXI          -1.1
    UPT.G5XI  UPM...XI        -.0073   UCW...XI        -.1131
    UPT.G5XI  CUT.DSXI         -.008
    B.FNXI    B...XI             -1.
    QVO0FHXI  BFH.VOXI           -1.   XGS.FHXI          -10.
    QWT0FHXI  BFH...XI            1.   BFH.WTXI           -1.
    QWT0FHXI  XSU.FHXI           -.3   XGS.FHXI            1.
    QVO0FLXI  BFL.VOXI           -1.   XGS.FLXI          -10.
    QWT0FLXI  BFL...XI            1.   BFL.WTXI           -1.
    QWT0FLXI  XSU.FLXI           -.3   XGS.FLXI            1.
    QVO0FNXI  BFN.VOXI           -1.   XGS.FNXI          -10.
    QWT0FNXI  BFN...XI            1.   BFN.WTXI           -1.
    QWT0FNXI  XSU.FNXI           -1.   XGS.FNXI            1.
    QVO0F1XI  BF1.VOXI           -1.   XGS.F1XI          -10.
    QWT0F1XI  BF1...XI            1.   BF1.WTXI           -1.
    QWT0F1XI  XSU.F1XI           -.3   XGS.F1XI            1.
    QVO0F2XI  BF2.VOXI           -1.   XGS.F2XI          -10.
    QWT0F2XI  BF2...XI            1.   BF2.WTXI           -1.
    QWT0F2XI  XSU.F2XI          -10.   XGS.F2XI            1.
    QVO0F3XI  BF3.VOXI           -1.   XGS.F3XI          -10.
    QWT0F3XI  BF3...XI            1.   BF3.WTXI           -1.
    QWT0F3XI  XSU.F3XI          -10.   XGS.F3XI            1.
    QWT0F3XI  XSL.F3XI          -.15
    QVO0F4XI  BF4.VOXI           -1.   XGS.F4XI          -10.
    QWT0F4XI  BF4...XI            1.   BF4.WTXI           -1.
    QWT0F4XI  XSU.F4XI           -.3   XGS.F4XI            1.
RHS
BOUNDS
 UP 0.BOUND   DFH...XI          100.
 UP 0.BOUND   DFN...XI          100.
 UP 0.BOUND   DF1...XI          100.
 UP 0.BOUND   DF2...XI          100.
 UP 0.BOUND   DF3...XI          100.
 UP 0.BOUND   DF4...XI          100.
 UP 0.BOUND   DL2...XI          100.
 UP 0.BOUND   DL3...XI          100.
 UP 0.BOUND   DL4...XI          100.
 UP 0.BOUND   DMM...XI          100.
 UP 0.BOUND   DPS...XI          100.
 FX 0.BOUND   EMR...XI       17.9327
 LO 0.BOUND   KLQ.PRXI           10.
ENDATA
