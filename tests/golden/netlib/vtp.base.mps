NAME          VTP.BASE
ROWS
 E  FIC.....
 E  FIP.....
 E  FOC.....
 E  FPC.....
 E  FRP.....
 E  FSC.....
 N  FAT...J.
 E  B...G1TA
 E  B...G2TA
 E  B...G3TA
 E  B...G4TA
 E  LY....TA
 E  T.....TA
 E  B...G1TB
 E  B...G2TB
 E  B...G3TB
 E  B...G4TB
 E  LY....TB
 E  T.....TB
 E  B...G1TC
 E  B...G2TC
 E  B...G3TC
 E  B...G4TC
 E  LY....TC
 E  T.....TC
 E  B...G1TD
 E  B...G2TD
 E  B...G3TD
 E  B...G4TD
 E  LY....TD
 E  T.....TD
 E  B...G1TE
 E  B...G2TE
 E  B...G3TE
 E  B...G4TE
 E  LY....TE
 E  T.....TE
 E  B...G1TF
 E  B...G2TF
 E  B...G3TF
 E  B...G4TF
 E  LY....TF
 E  T.....TF
 E  B...G1TG
 E  B...G2TG
 E  B...G3TG
 E  B...G4TG
 E  LY....TG
 E  T.....TG
 E  B...G1TH
 E  B...G2TH
 E  B...G3TH
 E  B...G4TH
 E  LY....TH
 E  T.....TH
 E  LBL1....
 L  LMP1..TA
 L  LNP1..TA
 L  LMP1..TB
 L  LNP1..TB
 L  R.P1..TB
 L  LMP1..TC
 L  LNP1..TC
 L  LBL1MNTC
 L  LBL1MXTC
 L  R.P1..TC
 L  LMP1..TD
 L  LNP1..TD
 L  LBL1MNTD
 L  LBL1MXTD
 L  R.P1..TD
 L  LMP1..TE
 L  LNP1..TE
 L  LBL1MNTE
 L  LBL1MXTE
 L  R.P1..TE
 L  LMP1..TF
 L  LNP1..TF
 L  LBL1MNTF
 L  LBL1MXTF
 L  R.P1..TF
 L  LMP1..TG
 L  LNP1..TG
 L  LBL1MNTG
 L  LBL1MXTG
 L  R.P1..TG
 G  BS.12.TH
 G  BS.123TH
 G  BS.124TH
 G  BS.13.TH
 G  BS.134TH
 G  BS.14.TH
 L  LMP1..TH
 L  LNP1..TH
 L  R.P1..TH
 L  LMP2..TA
 L  LNP2..TA
 L  LMP2..TB
 L  LNP2..TB
 L  R.P2..TB
 L  LMP2..TC
 L  LNP2..TC
 L  R.P2..TC
 L  LMP2..TD
 L  LNP2..TD
 L  R.P2..TD
 L  LMP2..TE
 L  LNP2..TE
 L  R.P2..TE
 L  LMP2..TF
 L  LNP2..TF
 L  R.P2..TF
 L  LMP2..TG
 L  LNP2..TG
 L  R.P2..TG
 G  BS.23.TH
 G  BS.234TH
 G  BS.24.TH
 L  LMP2..TH
 L  LNP2..TH
 L  R.P2..TH
 L  LMP3..TA
 L  LNP3..TA
 L  LSP3..TA
 L  LMP3..TB
 L  LNP3..TB
 L  LSP3..TB
 L  R.P3..TB
 L  LMP3..TC
 L  LNP3..TC
 L  LSP3..TC
 L  R.P3..TC
 L  LMP3..TD
 L  LNP3..TD
 L  LSP3..TD
 L  R.P3..TD
 L  LMP3..TE
 L  LNP3..TE
 L  LSP3..TE
 L  R.P3..TE
 L  LMP3..TF
 L  LNP3..TF
 L  LSP3..TF
 L  R.P3..TF
 L  LMP3..TG
 L  LNP3..TG
 L  LSP3..TG
 L  R.P3..TG
 G  BS.34.TH
 L  LMP3..TH
 L  LNP3..TH
 L  LSP3..TH
 L  R.P3..TH
 L  LMP4..TA
 L  LNP4..TA
 L  LMP4..TB
 L  LNP4..TB
 L  R.P4..TB
 L  LMP4..TC
 L  LNP4..TC
 L  R.P4..TC
 L  LMP4..TD
 L  LNP4..TD
 L  R.P4..TD
 L  LMP4..TE
 L  LNP4..TE
 L  R.P4..TE
 L  LMP4..TF
 L  LNP4..TF
 L  R.P4..TF
 L  LMP4..TG
 L  LNP4..TG
 L  R.P4..TG
 L  LMP4..TH
 L  LNP4..TH
 L  R.P4..TH
 L  LMP5..TA
 L  LNP5..TA
 L  LMP5..TB
 L  LNP5..TB
 L  R.P5..TB
 L  LMP5..TC
 L  LNP5..TC
 L  R.P5..TC
 L  LMP5..TD
 L  LNP5..TD
 L  R.P5..TD
 L  LMP5..TE
 L  LNP5..TE
 L  R.P5..TE
 L  LMP5..TF
 L  LNP5..TF
 L  R.P5..TF
 L  LMP5..TG
 L  LNP5..TG
 L  R.P5..TG
 L  LMP5..TH
 L  LNP5..TH
 L  R.P5..TH
COLUMNS
    FIC.....  FIC.....           -1.   FAT...J.            1.
    FIP.....  FIP.....           -1.   FAT...J.            1.
    FOC.....  FOC.....           -1.   FAT...J.            1.
    FPC.....  FPC.....           -1.   FAT...J.            1.
    FRP.....  FRP.....           -1.   FAT...J.            1.
    FSC.....  FSC.....           -1.   FAT...J.            1.
    KL..TCL1  LBL1....            1.   LBL1MNTC          100.
    KL..TCL1  LBL1MXTC          180.   FIC.....    1109.58904
    KL..TCL1  B...G3TC         -300.
    KL..TDL1  LBL1....            1.   LBL1MNTD          100.
    KL..TDL1  LBL1MXTD          180.   FIC.....    1109.58904
    KL..TDL1  B...G3TD         -300.
    KL..TEL1  LBL1....            1.   LBL1MNTE          100.
    KL..TEL1  LBL1MXTE          180.   FIC.....    1109.58904
    KL..TEL1  B...G3TE         -300.
    KL..TFL1  LBL1....            1.   LBL1MNTF          100.
    KL..TFL1  LBL1MXTF          180.   FIC.....    1109.58904
    KL..TFL1  B...G3TF         -300.
    KL..TGL1  LBL1....            1.   LBL1MNTG          100.
    KL..TGL1  LBL1MXTG          180.   FIC.....    1109.58904
    KL..TGL1  B...G3TG         -300.
    INV.G1T.  B...G1TA            1.
    INV.G2T.  B...G2TA            1.
    INV.G3T.  B...G3TA            1.
    INV.G4T.  B...G4TA            1.
    INV.G1TA  B...G1TA           -1.   FIC.....       3.69863
    INV.G1TA  B...G1TB            1.
    INV.G2TA  B...G2TA           -1.   FIC.....       3.69863
    INV.G2TA  B...G2TB            1.
    INV.G3TA  B...G3TA           -1.   FIC.....       3.69863
    INV.G3TA  B...G3TB            1.
    INV.G4TA  B...G4TA           -1.   FIC.....       3.69863
    INV.G4TA  B...G4TB            1.
    INV.G1TB  B...G1TB           -1.   FIC.....       3.69863
    INV.G1TB  B...G1TC            1.
    INV.G2TB  B...G2TB           -1.   FIC.....       3.69863
    INV.G2TB  B...G2TC            1.
    INV.G3TB  B...G3TB           -1.   FIC.....       3.69863
    INV.G3TB  B...G3TC            1.
    INV.G4TB  B...G4TB           -1.   FIC.....       3.69863
    INV.G4TB  B...G4TC            1.
    INV.G1TC  B...G1TC           -1.   FIC.....       3.69863
    INV.G1TC  B...G1TD            1.
    INV.G2TC  B...G2TC           -1.   FIC.....       3.69863
    INV.G2TC  B...G2TD            1.
    INV.G3TC  B...G3TC           -1.   FIC.....       3.69863
    INV.G3TC  B...G3TD            1.
    INV.G4TC  B...G4TC           -1.   FIC.....       3.69863
    INV.G4TC  B...G4TD            1.
    INV.G1TD  B...G1TD           -1.   FIC.....       3.69863
    INV.G1TD  B...G1TE            1.
    INV.G2TD  B...G2TD           -1.   FIC.....       3.69863
    INV.G2TD  B...G2TE            1.
    INV.G3TD  B...G3TD           -1.   FIC.....       3.69863
    INV.G3TD  B...G3TE            1.
    INV.G4TD  B...G4TD           -1.   FIC.....       3.69863
    INV.G4TD  B...G4TE            1.
    INV.G1TE  B...G1TE           -1.   FIC.....       3.69863
    INV.G1TE  B...G1TF            1.
    INV.G2TE  B...G2TE           -1.   FIC.....       3.69863
    INV.G2TE  B...G2TF            1.
    INV.G3TE  B...G3TE           -1.   FIC.....       3.69863
    INV.G3TE  B...G3TF            1.
    INV.G4TE  B...G4TE           -1.   FIC.....       3.69863
    INV.G4TE  B...G4TF            1.
    INV.G1TF  B...G1TF           -1.   FIC.....       3.69863
    INV.G1TF  B...G1TG            1.
    INV.G2TF  B...G2TF           -1.   FIC.....       3.69863
    INV.G2TF  B...G2TG            1.
    INV.G3TF  B...G3TF           -1.   FIC.....       3.69863
    INV.G3TF  B...G3TG            1.
    INV.G4TF  B...G4TF           -1.   FIC.....       3.69863
    INV.G4TF  B...G4TG            1.
    INV.G1TG  B...G1TG           -1.   FIC.....       3.69863
    INV.G1TG  B...G1TH            1.
    INV.G2TG  B...G2TG           -1.   FIC.....       3.69863
    INV.G2TG  B...G2TH            1.
    INV.G3TG  B...G3TG           -1.   FIC.....       3.69863
    INV.G3TG  B...G3TH            1.
    INV.G4TG  B...G4TG           -1.   FIC.....       3.69863
    INV.G4TG  B...G4TH            1.
    INV.G1TH  B...G1TH           -1.   FIC.....       3.69863
    INV.G1TH  BS.12.TH            2.   BS.123TH            .5
    INV.G1TH  BS.124TH            .5   BS.13.TH        .66667
    INV.G1TH  BS.134TH        .33333   BS.14.TH        .66667
    INV.G2TH  B...G2TH           -1.   FIC.....       3.69863
    INV.G2TH  BS.12.TH        .51282   BS.123TH        .25641
    INV.G2TH  BS.124TH        .25641   BS.23.TH        .51282
    INV.G2TH  BS.234TH        .25641   BS.24.TH        .51282
    INV.G3TH  B...G3TH           -1.   FIC.....       3.69863
    INV.G3TH  BS.123TH            .2   BS.13.TH        .26667
    INV.G3TH  BS.134TH        .13333   BS.23.TH            .8
    INV.G3TH  BS.234TH            .2   BS.34.TH        .26667
    INV.G4TH  B...G4TH           -1.   FIC.....       3.69863
    INV.G4TH  BS.124TH        .41667   BS.134TH        .27778
    INV.G4TH  BS.14.TH        .55556   BS.234TH        .41667
    INV.G4TH  BS.24.TH       1.66667   BS.34.TH        .55556
    INVAG1TA  B...G1TA           -1.   FIP.....         -100.
    INVAG1TA  B...G1TB            1.
    INVAG2TA  B...G2TA           -1.   FIP.....         -100.
    INVAG2TA  B...G2TB            1.
    INVAG3TA  B...G3TA           -1.   FIP.....         -100.
    INVAG3TA  B...G3TB            1.
    INVAG4TA  B...G4TA           -1.   FIP.....         -100.
    INVAG4TA  B...G4TB            1.
    INVAG1TB  B...G1TB           -1.   FIP.....         -87.5
    INVAG1TB  B...G1TC            1.
    INVAG2TB  B...G2TB           -1.   FIP.....         -87.5
    INVAG2TB  B...G2TC            1.
    INVAG3TB  B...G3TB           -1.   FIP.....         -87.5
    INVAG3TB  B...G3TC            1.
    INVAG4TB  B...G4TB           -1.   FIP.....         -87.5
    INVAG4TB  B...G4TC            1.
    INVAG1TC  B...G1TC           -1.   FIP.....          -75.
    INVAG1TC  B...G1TD            1.
    INVAG2TC  B...G2TC           -1.   FIP.....          -75.
    INVAG2TC  B...G2TD            1.
    INVAG3TC  B...G3TC           -1.   FIP.....          -75.
    INVAG3TC  B...G3TD            1.
    INVAG4TC  B...G4TC           -1.   FIP.....          -75.
    INVAG4TC  B...G4TD            1.
    INVAG1TD  B...G1TD           -1.   FIP.....         -62.5
    INVAG1TD  B...G1TE            1.
    INVAG2TD  B...G2TD           -1.   FIP.....         -62.5
    INVAG2TD  B...G2TE            1.
    INVAG3TD  B...G3TD           -1.   FIP.....         -62.5
    INVAG3TD  B...G3TE            1.
    INVAG4TD  B...G4TD           -1.   FIP.....         -62.5
    INVAG4TD  B...G4TE            1.
    INVAG1TE  B...G1TE           -1.   FIP.....          -50.
    INVAG1TE  B...G1TF            1.
    INVAG2TE  B...G2TE           -1.   FIP.....          -50.
    INVAG2TE  B...G2TF            1.
    INVAG3TE  B...G3TE           -1.   FIP.....          -50.
    INVAG3TE  B...G3TF            1.
    INVAG4TE  B...G4TE           -1.   FIP.....          -50.
    INVAG4TE  B...G4TF            1.
    INVAG1TF  B...G1TF           -1.   FIP.....         -37.5
    INVAG1TF  B...G1TG            1.
    INVAG2TF  B...G2TF           -1.   FIP.....         -37.5
    INVAG2TF  B...G2TG            1.
    INVAG3TF  B...G3TF           -1.   FIP.....         -37.5
    INVAG3TF  B...G3TG            1.
    INVAG4TF  B...G4TF           -1.   FIP.....         -37.5
    INVAG4TF  B...G4TG            1.
    INVAG1TG  B...G1TG           -1.   FIP.....          -25.
    INVAG1TG  B...G1TH            1.
    INVAG2TG  B...G2TG           -1.   FIP.....          -25.
    INVAG2TG  B...G2TH            1.
    INVAG3TG  B...G3TG           -1.   FIP.....          -25.
    INVAG3TG  B...G3TH            1.
    INVAG4TG  B...G4TG           -1.   FIP.....          -25.
    INVAG4TG  B...G4TH            1.
    INVAG1TH  B...G1TH           -1.   FIP.....         -12.5
    INVAG2TH  B...G2TH           -1.   FIP.....         -12.5
    INVAG3TH  B...G3TH           -1.   FIP.....         -12.5
    INVAG4TH  B...G4TH           -1.   FIP.....         -12.5
    TIME..T.  B...G1TA           10.   B...G2TA           13.
    TIME..T.  B...G3TA           25.   B...G4TA           12.
    TIME..T.  T.....TA            1.
    TIME..TA  B...G1TA          -10.   B...G2TA          -13.
    TIME..TA  B...G3TA          -25.   B...G4TA          -12.
    TIME..TA  T.....TA           -1.   B...G1TB           10.
    TIME..TA  B...G2TB           13.   B...G3TB           25.
    TIME..TA  B...G4TB           12.   T.....TB            1.
    TIME..TB  B...G1TB          -10.   B...G2TB          -13.
    TIME..TB  B...G3TB          -25.   B...G4TB          -12.
    TIME..TB  T.....TB           -1.   B...G1TC           10.
    TIME..TB  B...G2TC           13.   B...G3TC           25.
    TIME..TB  B...G4TC           12.   T.....TC            1.
    TIME..TB  LBL1MXTC            1.
    TIME..TC  B...G1TC          -10.   B...G2TC          -13.
    TIME..TC  B...G3TC          -25.   B...G4TC          -12.
    TIME..TC  T.....TC           -1.   B...G1TD           10.
    TIME..TC  B...G2TD           13.   B...G3TD           25.
    TIME..TC  B...G4TD           12.   T.....TD            1.
    TIME..TC  LBL1MNTC           -1.   LBL1MXTD            1.
    TIME..TD  B...G1TD          -10.   B...G2TD          -13.
    TIME..TD  B...G3TD          -25.   B...G4TD          -12.
    TIME..TD  T.....TD           -1.   B...G1TE           10.
    TIME..TD  B...G2TE           13.   B...G3TE           25.
    TIME..TD  B...G4TE           12.   T.....TE            1.
    TIME..TD  LBL1MNTD           -1.   LBL1MXTE            1.
    TIME..TE  B...G1TE          -10.   B...G2TE          -13.
    TIME..TE  B...G3TE          -25.   B...G4TE          -12.
    TIME..TE  T.....TE           -1.   B...G1TF           10.
    TIME..TE  B...G2TF           13.   B...G3TF           25.
    TIME..TE  B...G4TF           12.   T.....TF            1.
    TIME..TE  LBL1MNTE           -1.   LBL1MXTF            1.
    TIME..TF  B...G1TF          -10.   B...G2TF          -13.
    TIME..TF  B...G3TF          -25.   B...G4TF          -12.
    TIME..TF  T.....TF           -1.   B...G1TG           10.
    TIME..TF  B...G2TG           13.   B...G3TG           25.
    TIME..TF  B...G4TG           12.   T.....TG            1.
    TIME..TF  LBL1MNTF           -1.   LBL1MXTG            1.
    TIME..TG  B...G1TG          -10.   B...G2TG          -13.
    TIME..TG  B...G3TG          -25.   B...G4TG          -12.
    TIME..TG  T.....TG           -1.   B...G1TH           10.
    TIME..TG  B...G2TH           13.   B...G3TH           25.
    TIME..TG  B...G4TH           12.   T.....TH            1.
    TIME..TG  LBL1MNTG           -1.
    TIME..TH  B...G1TH          -10.   B...G2TH          -13.
    TIME..TH  B...G3TH          -25.   B...G4TH          -12.
    TIME..TH  T.....TH           -1.
    KYP1..TA  B...G1TA          -50.   LMP1..TA          -60.
    KYP1..TA  LNP1..TA           15.   LY....TA            1.
    KYP1..TA  R.P2..TB         4000.   R.P3..TB         3000.
    KYP1..TA  R.P4..TB         4000.   R.P5..TB         2500.
    YTP1..TA  B...G1TA           80.   LMP1..TA            1.
    YTP1..TA  LNP1..TA           -1.   T.....TA            1.
    YTP1..TA  FOC.....           60.
    KYP1..TB  B...G1TB          -50.   LMP1..TB          -60.
    KYP1..TB  LNP1..TB           15.   LY....TB            1.
    KYP1..TB  R.P1..TB         4000.   R.P2..TC         4000.
    KYP1..TB  R.P3..TC         3000.   R.P4..TC         4000.
    KYP1..TB  R.P5..TC         2500.
    P.P1..TB  R.P1..TB           -1.   FSC.....            1.
    YTP1..TB  B...G1TB           80.   LMP1..TB            1.
    YTP1..TB  LNP1..TB           -1.   T.....TB            1.
    YTP1..TB  FOC.....           60.
    KYP1..TC  B...G1TC          -50.   LMP1..TC          -60.
    KYP1..TC  LNP1..TC           15.   LY....TC            1.
    KYP1..TC  R.P1..TC         4000.   R.P2..TD         4000.
    KYP1..TC  R.P3..TD         3000.   R.P4..TD         4000.
    KYP1..TC  R.P5..TD         2500.
    P.P1..TC  R.P1..TC           -1.   FSC.....            1.
    YTP1..TC  B...G1TC           80.   LMP1..TC            1.
    YTP1..TC  LNP1..TC           -1.   T.....TC            1.
    YTP1..TC  FOC.....           60.
    KYP1..TD  B...G1TD          -50.   LMP1..TD          -60.
    KYP1..TD  LNP1..TD           15.   LY....TD            1.
    KYP1..TD  R.P1..TD         4000.   R.P2..TE         4000.
    KYP1..TD  R.P3..TE         3000.   R.P4..TE         4000.
    KYP1..TD  R.P5..TE         2500.
    P.P1..TD  R.P1..TD           -1.   FSC.....            1.
    YTP1..TD  B...G1TD           80.   LMP1..TD            1.
    YTP1..TD  LNP1..TD           -1.   T.....TD            1.
    YTP1..TD  FOC.....           60.
    KYP1..TE  B...G1TE          -50.   LMP1..TE          -60.
    KYP1..TE  LNP1..TE           15.   LY....TE            1.
    KYP1..TE  R.P1..TE         4000.   R.P2..TF         4000.
    KYP1..TE  R.P3..TF         3000.   R.P4..TF         4000.
    KYP1..TE  R.P5..TF         2500.
    P.P1..TE  R.P1..TE           -1.   FSC.....            1.
    YTP1..TE  B...G1TE           80.   LMP1..TE            1.
    YTP1..TE  LNP1..TE           -1.   T.....TE            1.
    YTP1..TE  FOC.....           60.
    KYP1..TF  B...G1TF          -50.   LMP1..TF          -60.
    KYP1..TF  LNP1..TF           15.   LY....TF            1.
    KYP1..TF  R.P1..TF         4000.   R.P2..TG         4000.
    KYP1..TF  R.P3..TG         3000.   R.P4..TG         4000.
    KYP1..TF  R.P5..TG         2500.
    P.P1..TF  R.P1..TF           -1.   FSC.....            1.
    YTP1..TF  B...G1TF           80.   LMP1..TF            1.
    YTP1..TF  LNP1..TF           -1.   T.....TF            1.
    YTP1..TF  FOC.....           60.
    KYP1..TG  B...G1TG          -50.   LMP1..TG          -60.
    KYP1..TG  LNP1..TG           15.   LY....TG            1.
    KYP1..TG  R.P1..TG         4000.   R.P2..TH         4000.
    KYP1..TG  R.P3..TH         3000.   R.P4..TH         4000.
    KYP1..TG  R.P5..TH         2500.
    P.P1..TG  R.P1..TG           -1.   FSC.....            1.
    YTP1..TG  B...G1TG           80.   LMP1..TG            1.
    YTP1..TG  LNP1..TG           -1.   T.....TG            1.
    YTP1..TG  FOC.....           60.
    KYP1..TH  B...G1TH          -50.   LMP1..TH          -60.
    KYP1..TH  LNP1..TH           15.   LY....TH            1.
    KYP1..TH  R.P1..TH         4000.
    P.P1..TH  R.P1..TH           -1.   FSC.....            1.
    YTP1..TH  B...G1TH           80.   LMP1..TH            1.
    YTP1..TH  LNP1..TH           -1.   T.....TH            1.
    YTP1..TH  FOC.....           60.
    KYP2..TA  B...G2TA          -20.   LMP2..TA          -60.
    KYP2..TA  LNP2..TA            5.   LY....TA            1.
    KYP2..TA  LSP3..TB           -1.   R.P1..TB         4000.
    KYP2..TA  R.P3..TB         4000.   R.P4..TB         3000.
    KYP2..TA  R.P5..TB         2000.
    YTP2..TA  B...G2TA           93.   LMP2..TA            1.
    YTP2..TA  LNP2..TA           -1.   T.....TA            1.
    YTP2..TA  FOC.....           30.
    KYP2..TB  B...G2TB          -20.   LMP2..TB          -60.
    KYP2..TB  LNP2..TB            5.   LY....TB            1.
    KYP2..TB  LSP3..TA           -1.   R.P2..TB         4000.
    KYP2..TB  LSP3..TC           -1.   R.P1..TC         4000.
    KYP2..TB  R.P3..TC         4000.   R.P4..TC         3000.
    KYP2..TB  R.P5..TC         2000.
    P.P2..TB  R.P2..TB           -1.   FSC.....            1.
    YTP2..TB  B...G2TB           93.   LMP2..TB            1.
    YTP2..TB  LNP2..TB           -1.   T.....TB            1.
    YTP2..TB  FOC.....           30.
    KYP2..TC  B...G2TC          -20.   LMP2..TC          -60.
    KYP2..TC  LNP2..TC            5.   LY....TC            1.
    KYP2..TC  LSP3..TB           -1.   R.P2..TC         4000.
    KYP2..TC  LSP3..TD           -1.   R.P1..TD         4000.
    KYP2..TC  R.P3..TD         4000.   R.P4..TD         3000.
    KYP2..TC  R.P5..TD         2000.
    P.P2..TC  R.P2..TC           -1.   FSC.....            1.
    YTP2..TC  B...G2TC           93.   LMP2..TC            1.
    YTP2..TC  LNP2..TC           -1.   T.....TC            1.
    YTP2..TC  FOC.....           30.
    KYP2..TD  B...G2TD          -20.   LMP2..TD          -60.
    KYP2..TD  LNP2..TD            5.   LY....TD            1.
    KYP2..TD  LSP3..TC           -1.   R.P2..TD         4000.
    KYP2..TD  LSP3..TE           -1.   R.P1..TE         4000.
    KYP2..TD  R.P3..TE         4000.   R.P4..TE         3000.
    KYP2..TD  R.P5..TE         2000.
    P.P2..TD  R.P2..TD           -1.   FSC.....            1.
    YTP2..TD  B...G2TD           93.   LMP2..TD            1.
    YTP2..TD  LNP2..TD           -1.   T.....TD            1.
    YTP2..TD  FOC.....           30.
    KYP2..TE  B...G2TE          -20.   LMP2..TE          -60.
    KYP2..TE  LNP2..TE            5.   LY....TE            1.
    KYP2..TE  LSP3..TD           -1.   R.P2..TE         4000.
    KYP2..TE  LSP3..TF           -1.   R.P1..TF         4000.
    KYP2..TE  R.P3..TF         4000.   R.P4..TF         3000.
    KYP2..TE  R.P5..TF         2000.
    P.P2..TE  R.P2..TE           -1.   FSC.....            1.
    YTP2..TE  B...G2TE           93.   LMP2..TE            1.
    YTP2..TE  LNP2..TE           -1.   T.....TE            1.
    YTP2..TE  FOC.....           30.
    KYP2..TF  B...G2TF          -20.   LMP2..TF          -60.
    KYP2..TF  LNP2..TF            5.   LY....TF            1.
    KYP2..TF  LSP3..TE           -1.   R.P2..TF         4000.
    KYP2..TF  LSP3..TG           -1.   R.P1..TG         4000.
    KYP2..TF  R.P3..TG         4000.   R.P4..TG         3000.
    KYP2..TF  R.P5..TG         2000.
    P.P2..TF  R.P2..TF           -1.   FSC.....            1.
    YTP2..TF  B...G2TF           93.   LMP2..TF            1.
    YTP2..TF  LNP2..TF           -1.   T.....TF            1.
    YTP2..TF  FOC.....           30.
    KYP2..TG  B...G2TG          -20.   LMP2..TG          -60.
    KYP2..TG  LNP2..TG            5.   LY....TG            1.
    KYP2..TG  LSP3..TF           -1.   R.P2..TG         4000.
    KYP2..TG  LSP3..TH           -1.   R.P1..TH         4000.
    KYP2..TG  R.P3..TH         4000.   R.P4..TH         3000.
    KYP2..TG  R.P5..TH         2000.
    P.P2..TG  R.P2..TG           -1.   FSC.....            1.
    YTP2..TG  B...G2TG           93.   LMP2..TG            1.
    YTP2..TG  LNP2..TG           -1.   T.....TG            1.
    YTP2..TG  FOC.....           30.
    KYP2..TH  B...G2TH          -20.   LMP2..TH          -60.
    KYP2..TH  LNP2..TH            5.   LY....TH            1.
    KYP2..TH  LSP3..TG           -1.   R.P2..TH         4000.
    P.P2..TH  R.P2..TH           -1.   FSC.....            1.
    YTP2..TH  B...G2TH           93.   LMP2..TH            1.
    YTP2..TH  LNP2..TH           -1.   T.....TH            1.
    YTP2..TH  FOC.....           30.
    KYP3..TA  B...G3TA          -50.   LMP3..TA          -60.
    KYP3..TA  LNP3..TA           15.   LY....TA            1.
    KYP3..TA  LSP3..TA            1.   R.P1..TB         4000.
    KYP3..TA  R.P2..TB         3000.   R.P4..TB         4000.
    KYP3..TA  R.P5..TB         2500.
    YTP3..TA  B...G3TA           65.   LMP3..TA            1.
    YTP3..TA  LNP3..TA           -1.   T.....TA            1.
    YTP3..TA  FOC.....           40.
    KYP3..TB  B...G3TB          -50.   LMP3..TB          -60.
    KYP3..TB  LNP3..TB           15.   LY....TB            1.
    KYP3..TB  LSP3..TB            1.   R.P3..TB         4000.
    KYP3..TB  R.P1..TC         4000.   R.P2..TC         3000.
    KYP3..TB  R.P4..TC         4000.   R.P5..TC         2500.
    P.P3..TB  R.P3..TB           -1.   FSC.....            1.
    YTP3..TB  B...G3TB           65.   LMP3..TB            1.
    YTP3..TB  LNP3..TB           -1.   T.....TB            1.
    YTP3..TB  FOC.....           40.
    KYP3..TC  B...G3TC          -50.   LMP3..TC          -60.
    KYP3..TC  LNP3..TC           15.   LY....TC            1.
    KYP3..TC  LSP3..TC            1.   R.P3..TC         4000.
    KYP3..TC  R.P1..TD         4000.   R.P2..TD         3000.
    KYP3..TC  R.P4..TD         4000.   R.P5..TD         2500.
    P.P3..TC  R.P3..TC           -1.   FSC.....            1.
    YTP3..TC  B...G3TC           65.   LMP3..TC            1.
    YTP3..TC  LNP3..TC           -1.   T.....TC            1.
    YTP3..TC  FOC.....           40.
    KYP3..TD  B...G3TD          -50.   LMP3..TD          -60.
    KYP3..TD  LNP3..TD           15.   LY....TD            1.
    KYP3..TD  LSP3..TD            1.   R.P3..TD         4000.
    KYP3..TD  R.P1..TE         4000.   R.P2..TE         3000.
    KYP3..TD  R.P4..TE         4000.   R.P5..TE         2500.
    P.P3..TD  R.P3..TD           -1.   FSC.....            1.
    YTP3..TD  B...G3TD           65.   LMP3..TD            1.
    YTP3..TD  LNP3..TD           -1.   T.....TD            1.
    YTP3..TD  FOC.....           40.
    KYP3..TE  B...G3TE          -50.   LMP3..TE          -60.
    KYP3..TE  LNP3..TE           15.   LY....TE            1.
    KYP3..TE  LSP3..TE            1.   R.P3..TE         4000.
    KYP3..TE  R.P1..TF         4000.   R.P2..TF         3000.
    KYP3..TE  R.P4..TF         4000.   R.P5..TF         2500.
    P.P3..TE  R.P3..TE           -1.   FSC.....            1.
    YTP3..TE  B...G3TE           65.   LMP3..TE            1.
    YTP3..TE  LNP3..TE           -1.   T.....TE            1.
    YTP3..TE  FOC.....           40.
    KYP3..TF  B...G3TF          -50.   LMP3..TF          -60.
    KYP3..TF  LNP3..TF           15.   LY....TF            1.
    KYP3..TF  LSP3..TF            1.   R.P3..TF         4000.
    KYP3..TF  R.P1..TG         4000.   R.P2..TG         3000.
    KYP3..TF  R.P4..TG         4000.   R.P5..TG         2500.
    P.P3..TF  R.P3..TF           -1.   FSC.....            1.
    YTP3..TF  B...G3TF           65.   LMP3..TF            1.
    YTP3..TF  LNP3..TF           -1.   T.....TF            1.
    YTP3..TF  FOC.....           40.
    KYP3..TG  B...G3TG          -50.   LMP3..TG          -60.
    KYP3..TG  LNP3..TG           15.   LY....TG            1.
    KYP3..TG  LSP3..TG            1.   R.P3..TG         4000.
    KYP3..TG  R.P1..TH         4000.   R.P2..TH         3000.
    KYP3..TG  R.P4..TH         4000.   R.P5..TH         2500.
    P.P3..TG  R.P3..TG           -1.   FSC.....            1.
    YTP3..TG  B...G3TG           65.   LMP3..TG            1.
    YTP3..TG  LNP3..TG           -1.   T.....TG            1.
    YTP3..TG  FOC.....           40.
    KYP3..TH  B...G3TH          -50.   LMP3..TH          -60.
    KYP3..TH  LNP3..TH           15.   LY....TH            1.
    KYP3..TH  LSP3..TH            1.   R.P3..TH         4000.
    P.P3..TH  R.P3..TH           -1.   FSC.....            1.
    YTP3..TH  B...G3TH           65.   LMP3..TH            1.
    YTP3..TH  LNP3..TH           -1.   T.....TH            1.
    YTP3..TH  FOC.....           40.
    KYP4..TA  B...G4TA          -50.   LMP4..TA          -60.
    KYP4..TA  LNP4..TA           15.   LY....TA            1.
    KYP4..TA  R.P1..TB         3000.   R.P2..TB         4000.
    KYP4..TA  R.P3..TB         4000.   R.P5..TB         2500.
    YTP4..TA  B...G4TA           92.   LMP4..TA            1.
    YTP4..TA  LNP4..TA           -1.   T.....TA            1.
    YTP4..TA  FOC.....           50.
    KYP4..TB  B...G4TB          -50.   LMP4..TB          -60.
    KYP4..TB  LNP4..TB           15.   LY....TB            1.
    KYP4..TB  R.P4..TB         4000.   R.P1..TC         3000.
    KYP4..TB  R.P2..TC         4000.   R.P3..TC         4000.
    KYP4..TB  R.P5..TC         2500.
    P.P4..TB  R.P4..TB           -1.   FSC.....            1.
    YTP4..TB  B...G4TB           92.   LMP4..TB            1.
    YTP4..TB  LNP4..TB           -1.   T.....TB            1.
    YTP4..TB  FOC.....           50.
    KYP4..TC  B...G4TC          -50.   LMP4..TC          -60.
    KYP4..TC  LNP4..TC           15.   LY....TC            1.
    KYP4..TC  R.P4..TC         4000.   R.P1..TD         3000.
    KYP4..TC  R.P2..TD         4000.   R.P3..TD         4000.
    KYP4..TC  R.P5..TD         2500.
    P.P4..TC  R.P4..TC           -1.   FSC.....            1.
    YTP4..TC  B...G4TC           92.   LMP4..TC            1.
    YTP4..TC  LNP4..TC           -1.   T.....TC            1.
    YTP4..TC  FOC.....           50.
    KYP4..TD  B...G4TD          -50.   LMP4..TD          -60.
    KYP4..TD  LNP4..TD           15.   LY....TD            1.
    KYP4..TD  R.P4..TD         4000.   R.P1..TE         3000.
    KYP4..TD  R.P2..TE         4000.   R.P3..TE         4000.
    KYP4..TD  R.P5..TE         2500.
    P.P4..TD  R.P4..TD           -1.   FSC.....            1.
    YTP4..TD  B...G4TD           92.   LMP4..TD            1.
    YTP4..TD  LNP4..TD           -1.   T.....TD            1.
    YTP4..TD  FOC.....           50.
    KYP4..TE  B...G4TE          -50.   LMP4..TE          -60.
    KYP4..TE  LNP4..TE           15.   LY....TE            1.
    KYP4..TE  R.P4..TE         4000.   R.P1..TF         3000.
    KYP4..TE  R.P2..TF         4000.   R.P3..TF         4000.
    KYP4..TE  R.P5..TF         2500.
    P.P4..TE  R.P4..TE           -1.   FSC.....            1.
    YTP4..TE  B...G4TE           92.   LMP4..TE            1.
    YTP4..TE  LNP4..TE           -1.   T.....TE            1.
    YTP4..TE  FOC.....           50.
    KYP4..TF  B...G4TF          -50.   LMP4..TF          -60.
    KYP4..TF  LNP4..TF           15.   LY....TF            1.
    KYP4..TF  R.P4..TF         4000.   R.P1..TG         3000.
    KYP4..TF  R.P2..TG         4000.   R.P3..TG         4000.
    KYP4..TF  R.P5..TG         2500.
    P.P4..TF  R.P4..TF           -1.   FSC.....            1.
    YTP4..TF  B...G4TF           92.   LMP4..TF            1.
    YTP4..TF  LNP4..TF           -1.   T.....TF            1.
    YTP4..TF  FOC.....           50.
    KYP4..TG  B...G4TG          -50.   LMP4..TG          -60.
    KYP4..TG  LNP4..TG           15.   LY....TG            1.
    KYP4..TG  R.P4..TG         4000.   R.P1..TH         3000.
    KYP4..TG  R.P2..TH         4000.   R.P3..TH         4000.
    KYP4..TG  R.P5..TH         2500.
    P.P4..TG  R.P4..TG           -1.   FSC.....            1.
    YTP4..TG  B...G4TG           92.   LMP4..TG            1.
    YTP4..TG  LNP4..TG           -1.   T.....TG            1.
    YTP4..TG  FOC.....           50.
    KYP4..TH  B...G4TH          -50.   LMP4..TH          -60.
    KYP4..TH  LNP4..TH           15.   LY....TH            1.
    KYP4..TH  R.P4..TH         4000.
    P.P4..TH  R.P4..TH           -1.   FSC.....            1.
    YTP4..TH  B...G4TH           92.   LMP4..TH            1.
    YTP4..TH  LNP4..TH           -1.   T.....TH            1.
    YTP4..TH  FOC.....           50.
    KYP5..TA  LMP5..TA          -60.   LNP5..TA           20.
    KYP5..TA  LY....TA            1.
    YTP5..TA  LMP5..TA            1.   LNP5..TA           -1.
    YTP5..TA  T.....TA            1.
    KYP5..TB  LMP5..TB          -60.   LNP5..TB           20.
    KYP5..TB  LY....TB            1.   R.P5..TB         4000.
    P.P5..TB  R.P5..TB           -1.   FSC.....            1.
    YTP5..TB  LMP5..TB            1.   LNP5..TB           -1.
    YTP5..TB  T.....TB            1.
    KYP5..TC  LMP5..TC          -60.   LNP5..TC           20.
    KYP5..TC  LY....TC            1.   R.P5..TC         4000.
    P.P5..TC  R.P5..TC           -1.   FSC.....            1.
    YTP5..TC  LMP5..TC            1.   LNP5..TC           -1.
    YTP5..TC  T.....TC            1.
    KYP5..TD  LMP5..TD          -60.   LNP5..TD           20.
    KYP5..TD  LY....TD            1.   R.P5..TD         4000.
    P.P5..TD  R.P5..TD           -1.   FSC.....            1.
    YTP5..TD  LMP5..TD            1.   LNP5..TD           -1.
    YTP5..TD  T.....TD            1.
    KYP5..TE  LMP5..TE          -60.   LNP5..TE           20.
    KYP5..TE  LY....TE            1.   R.P5..TE         4000.
    P.P5..TE  R.P5..TE           -1.   FSC.....            1.
    YTP5..TE  LMP5..TE            1.   LNP5..TE           -1.
    YTP5..TE  T.....TE            1.
    KYP5..TF  LMP5..TF          -60.   LNP5..TF           20.
    KYP5..TF  LY....TF            1.   R.P5..TF         4000.
    P.P5..TF  R.P5..TF           -1.   FSC.....            1.
    YTP5..TF  LMP5..TF            1.   LNP5..TF           -1.
    YTP5..TF  T.....TF            1.
    KYP5..TG  LMP5..TG          -60.   LNP5..TG           20.
    KYP5..TG  LY....TG            1.   R.P5..TG         4000.
    P.P5..TG  R.P5..TG           -1.   FSC.....            1.
    YTP5..TG  LMP5..TG            1.   LNP5..TG           -1.
    YTP5..TG  T.....TG            1.
    KYP5..TH  LMP5..TH          -60.   LNP5..TH           20.
    KYP5..TH  LY....TH            1.   R.P5..TH         4000.
    P.P5..TH  R.P5..TH           -1.   FSC.....            1.
    YTP5..TH  LMP5..TH            1.   LNP5..TH           -1.
    YTP5..TH  T.....TH            1.
RHS
    RHS00001  LY....TA            1.   LY....TB            1.
    RHS00001  R.P1..TB         4000.   R.P2..TB         4000.
    RHS00001  R.P3..TB         4000.   R.P4..TB         4000.
    RHS00001  R.P5..TB         4000.   LY....TC            1.
    RHS00001  R.P1..TC         4000.   R.P2..TC         4000.
    RHS00001  R.P3..TC         4000.   R.P4..TC         4000.
    RHS00001  R.P5..TC         4000.   LY....TD            1.
    RHS00001  R.P1..TD         4000.   R.P2..TD         4000.
    RHS00001  R.P3..TD         4000.   R.P4..TD         4000.
    RHS00001  R.P5..TD         4000.   LY....TE            1.
    RHS00001  R.P1..TE         4000.   R.P2..TE         4000.
    RHS00001  R.P3..TE         4000.   R.P4..TE         4000.
    RHS00001  R.P5..TE         4000.   LY....TF            1.
    RHS00001  R.P1..TF         4000.   R.P2..TF         4000.
    RHS00001  R.P3..TF         4000.   R.P4..TF         4000.
    RHS00001  R.P5..TF         4000.   LY....TG            1.
    RHS00001  R.P1..TG         4000.   R.P2..TG         4000.
    RHS00001  R.P3..TG         4000.   R.P4..TG         4000.
    RHS00001  R.P5..TG         4000.   LY....TH            1.
    RHS00001  R.P1..TH         4000.   R.P2..TH         4000.
    RHS00001  R.P3..TH         4000.   R.P4..TH         4000.
    RHS00001  R.P5..TH         4000.   LBL1....            1.
    RHS00001  LBL1MXTC          280.   LBL1MXTD          280.
    RHS00001  LBL1MXTE          280.   LBL1MXTF          280.
    RHS00001  LBL1MXTG          280.   BS.12.TH     771.28205
    RHS00001  BS.123TH     315.64103   BS.124TH     337.30769
    RHS00001  BS.13.TH     346.66667   BS.134TH     261.11111
    RHS00001  BS.14.TH     375.55556   BS.23.TH     251.28205
    RHS00001  BS.234TH     207.30769   BS.24.TH     337.94872
    RHS00001  BS.34.TH     202.22222
BOUNDS
 FR BOUND     FOC.....
 UP BOUND     KL..TCL1            1.
 FX BOUND     KL..TDL1            0.
 FX BOUND     KL..TEL1            0.
 FX BOUND     KL..TFL1            0.
 FX BOUND     KL..TGL1            1.
 FX BOUND     INV.G1T.         1300.
 FX BOUND     INV.G2T.          600.
 FX BOUND     INV.G3T.          110.
 FX BOUND     INV.G4T.          900.
 LO BOUND     INV.G1TA          300.
 LO BOUND     INV.G2TA          100.
 LO BOUND     INV.G3TA          100.
 LO BOUND     INV.G4TA          100.
 LO BOUND     INV.G1TB          300.
 LO BOUND     INV.G2TB          100.
 LO BOUND     INV.G3TB          100.
 LO BOUND     INV.G4TB          100.
 LO BOUND     INV.G1TC          300.
 LO BOUND     INV.G2TC          100.
 LO BOUND     INV.G3TC          100.
 LO BOUND     INV.G4TC          100.
 LO BOUND     INV.G1TD          300.
 LO BOUND     INV.G2TD          100.
 LO BOUND     INV.G3TD          100.
 LO BOUND     INV.G4TD          100.
 LO BOUND     INV.G1TE          300.
 LO BOUND     INV.G2TE          100.
 LO BOUND     INV.G3TE          100.
 LO BOUND     INV.G4TE          100.
 LO BOUND     INV.G1TF          300.
 LO BOUND     INV.G2TF          100.
 LO BOUND     INV.G3TF          100.
 LO BOUND     INV.G4TF          100.
 LO BOUND     INV.G1TG          300.
 LO BOUND     INV.G2TG          100.
 LO BOUND     INV.G3TG          100.
 LO BOUND     INV.G4TG          100.
 LO BOUND     INV.G1TH          300.
 LO BOUND     INV.G2TH          100.
 LO BOUND     INV.G3TH          100.
 LO BOUND     INV.G4TH          100.
 LO BOUND     INVAG1TA         -300.
 UP BOUND     INVAG1TA            0.
 LO BOUND     INVAG2TA         -100.
 UP BOUND     INVAG2TA            0.
 LO BOUND     INVAG3TA         -100.
 UP BOUND     INVAG3TA            0.
 LO BOUND     INVAG4TA         -100.
 UP BOUND     INVAG4TA            0.
 LO BOUND     INVAG1TB         -300.
 UP BOUND     INVAG1TB            0.
 LO BOUND     INVAG2TB         -100.
 UP BOUND     INVAG2TB            0.
 LO BOUND     INVAG3TB         -100.
 UP BOUND     INVAG3TB            0.
 LO BOUND     INVAG4TB         -100.
 UP BOUND     INVAG4TB            0.
 LO BOUND     INVAG1TC         -300.
 UP BOUND     INVAG1TC            0.
 LO BOUND     INVAG2TC         -100.
 UP BOUND     INVAG2TC            0.
 LO BOUND     INVAG3TC         -100.
 UP BOUND     INVAG3TC            0.
 LO BOUND     INVAG4TC         -100.
 UP BOUND     INVAG4TC            0.
 LO BOUND     INVAG1TD         -300.
 UP BOUND     INVAG1TD            0.
 LO BOUND     INVAG2TD         -100.
 UP BOUND     INVAG2TD            0.
 LO BOUND     INVAG3TD         -100.
 UP BOUND     INVAG3TD            0.
 LO BOUND     INVAG4TD         -100.
 UP BOUND     INVAG4TD            0.
 LO BOUND     INVAG1TE         -300.
 UP BOUND     INVAG1TE            0.
 LO BOUND     INVAG2TE         -100.
 UP BOUND     INVAG2TE            0.
 LO BOUND     INVAG3TE         -100.
 UP BOUND     INVAG3TE            0.
 LO BOUND     INVAG4TE         -100.
 UP BOUND     INVAG4TE            0.
 LO BOUND     INVAG1TF         -300.
 UP BOUND     INVAG1TF            0.
 LO BOUND     INVAG2TF         -100.
 UP BOUND     INVAG2TF            0.
 LO BOUND     INVAG3TF         -100.
 UP BOUND     INVAG3TF            0.
 LO BOUND     INVAG4TF         -100.
 UP BOUND     INVAG4TF            0.
 LO BOUND     INVAG1TG         -300.
 UP BOUND     INVAG1TG            0.
 LO BOUND     INVAG2TG         -100.
 UP BOUND     INVAG2TG            0.
 LO BOUND     INVAG3TG         -100.
 UP BOUND     INVAG3TG            0.
 LO BOUND     INVAG4TG         -100.
 UP BOUND     INVAG4TG            0.
 LO BOUND     INVAG1TH         -300.
 UP BOUND     INVAG1TH            0.
 LO BOUND     INVAG2TH         -100.
 UP BOUND     INVAG2TH            0.
 LO BOUND     INVAG3TH         -100.
 UP BOUND     INVAG3TH            0.
 LO BOUND     INVAG4TH         -100.
 UP BOUND     INVAG4TH            0.
 FX BOUND     TIME..T.            0.
 FX BOUND     TIME..TH          180.
 UP BOUND     KYP1..TA            1.
 UP BOUND     KYP1..TB            1.
 UP BOUND     KYP1..TC            1.
 FX BOUND     KYP1..TD            1.
 UP BOUND     KYP1..TE            1.
 UP BOUND     KYP1..TF            1.
 UP BOUND     KYP1..TG            1.
 UP BOUND     KYP1..TH            1.
 UP BOUND     KYP2..TA            1.
 FX BOUND     KYP2..TB            1.
 UP BOUND     KYP2..TC            1.
 UP BOUND     KYP2..TD            1.
 UP BOUND     KYP2..TE            1.
 FX BOUND     KYP2..TF            1.
 UP BOUND     KYP2..TG            1.
 UP BOUND     KYP2..TH            1.
 FX BOUND     KYP3..TA            1.
 UP BOUND     KYP3..TB            1.
 UP BOUND     KYP3..TC            1.
 UP BOUND     KYP3..TD            1.
 FX BOUND     KYP3..TE            1.
 UP BOUND     KYP3..TF            1.
 FX BOUND     KYP3..TG            1.
 UP BOUND     KYP3..TH            1.
 UP BOUND     KYP4..TA            1.
 UP BOUND     KYP4..TB            1.
 FX BOUND     KYP4..TC            1.
 UP BOUND     KYP4..TD            1.
 UP BOUND     KYP4..TE            1.
 UP BOUND     KYP4..TF            1.
 UP BOUND     KYP4..TG            1.
 UP BOUND     KYP4..TH            1.
 UP BOUND     KYP5..TA            1.
 UP BOUND     KYP5..TB            1.
 UP BOUND     KYP5..TC            1.
 UP BOUND     KYP5..TD            1.
 UP BOUND     KYP5..TE            1.
 UP BOUND     KYP5..TF            1.
 UP BOUND     KYP5..TG            1.
 FX BOUND     KYP5..TH            1.
ENDATA
