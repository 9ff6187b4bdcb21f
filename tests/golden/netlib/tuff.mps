NAME          TUFF
ROWS
 N  B...ML..
 E  BA3...BW
 E  BA4...BW
 E  BA5...BW
 E  BBI...BW
 E  BCD...BW
 E  BCP...BW
 E  BC3...BW
 E  BC4...BW
 E  BFD...BW
 E  BFE...BW
 E  BFF...BW
 E  BFH...BW
 E  BFI...BW
 E  BFJ...BW
 E  BFK...BW
 E  BFL...BW
 E  BFM...BW
 E  BFN...BW
 E  BFO...BW
 E  BFP...BW
 E  BFQ...BW
 E  BFS...BW
 E  BFT...BW
 E  BFU...BW
 E  BFX...BW
 E  BFY...BW
 E  BF1...BW
 E  BF2...BW
 E  BF3...BW
 E  BGS...BW
 E  BG2...BW
 E  BG4...BW
 E  BG8...BW
 E  BH....BW
 E  BHC...BW
 E  BIS...BW
 E  BIT...BW
 E  BI4...BW
 E  BI5...BW
 E  BK1...BW
 E  BK2...BW
 E  BK3...BW
 E  BK4...BW
 E  BK5...BW
 E  BK6...BW
 E  BLF...BW
 E  BLH...BW
 E  BLI...BW
 E  BLL...BW
 E  BLP...BW
 E  BLV...BW
 E  BN4...BW
 E  BN5...BW
 E  BPE...BW
 E  BPG...BW
 E  BPH...BW
 E  BPI...BW
 E  BPQ...BW
 E  BPT...BW
 E  BPU...BW
 E  BPV...BW
 E  BPZ...BW
 E  BR1...BW
 E  BSA...BW
 E  BSB...BW
 E  BSC...BW
 E  BSD...BW
 E  BSE...BW
 E  BSF...BW
 E  BSG...BW
 E  BSH...BW
 E  BSI...BW
 E  BSJ...BW
 E  BSK...BW
 E  BSN...BW
 E  BSO...BW
 E  BSP...BW
 E  BSQ...BW
 E  BSR...BW
 E  BST...BW
 E  BSU...BW
 E  BTA...BW
 E  BTB...BW
 E  BTC...BW
 E  BTD...BW
 E  BTE...BW
 E  BTF...BW
 E  BTG...BW
 E  BTH...BW
 E  BTI...BW
 E  BTJ...BW
 E  BTK...BW
 E  BTL...BW
 E  BTM...BW
 E  BTN...BW
 E  BTO...BW
 E  BTP...BW
 E  BTQ...BW
 E  BTR...BW
 E  BTS...BW
 E  BTT...BW
 E  BTU...BW
 E  BTV...BW
 E  BTW...BW
 E  BTX...BW
 E  BTY...BW
 E  BTZ...BW
 E  BU4...BW
 E  BU5...BW
 E  BXA...BW
 E  BXB...BW
 E  BXC...BW
 E  BXD...BW
 E  BXE...BW
 E  BXF...BW
 E  BXG...BW
 E  BXH...BW
 E  BXI...BW
 E  BXJ...BW
 E  BXK...BW
 E  BXL...BW
 E  BXM...BW
 E  BXN...BW
 E  BXO...BW
 E  BXP...BW
 E  BXQ...BW
 E  BXR...BW
 E  BXS...BW
 E  BXT...BW
 E  BXU...BW
 E  BXV...BW
 E  BXW...BW
 E  BXX...BW
 E  BXZ...BW
 E  BX1...BW
 E  BX2...BW
 E  BX3...BW
 E  BX4...BW
 E  BX5...BW
 E  BX6...BW
 E  BYU...BW
 E  BZA...BW
 E  BZB...BW
 E  BZC...BW
 E  BZD...BW
 E  BZE...BW
 E  BZF...BW
 E  BZG...BW
 E  BZH...BW
 E  BZI...BW
 E  BZJ...BW
 E  B1A...BW
 E  B1B...BW
 E  B1C...BW
 E  B1D...BW
 E  B1G...BW
 E  B1H...BW
 E  B1I...BW
 E  B1J...BW
 E  B1K...BW
 E  B1L...BW
 E  B1M...BW
 E  B1N...BW
 E  B1O...BW
 E  B2A...BW
 E  B2B...BW
 E  B2C...BW
 E  B2D...BW
 E  B2G...BW
 E  B2H...BW
 E  B2I...BW
 E  B2J...BW
 E  B2K...BW
 E  B2L...BW
 E  B2M...BW
 E  B2N...BW
 E  B2O...BW
 E  B3A...BW
 E  B3B...BW
 E  B3C...BW
 E  B3D...BW
 E  B3G...BW
 E  B3H...BW
 E  B3I...BW
 E  B3J...BW
 E  B3K...BW
 E  B3L...BW
 E  B3M...BW
 E  B3N...BW
 E  B3O...BW
 E  B3U...BW
 E  B4....BW
 E  B4A...BW
 E  B4B...BW
 E  B4C...BW
 E  B4D...BW
 E  B4G...BW
 E  B4H...BW
 E  B4I...BW
 E  B4J...BW
 E  B4K...BW
 E  B4L...BW
 E  B4M...BW
 E  B4N...BW
 E  B4O...BW
 E  B4U...BW
 E  B4Z...BW
 E  B5A...BW
 E  B5B...BW
 E  B5C...BW
 E  B5D...BW
 E  B5G...BW
 E  B5H...BW
 E  B5I...BW
 E  B5J...BW
 E  B5K...BW
 E  B5L...BW
 E  B5M...BW
 E  B5N...BW
 E  B5O...BW
 E  B5S...BW
 E  B5U...BW
 E  B6....BW
 E  B8....BW
 E  B9....BW
 E  BG2.VOBW
 E  BG4.VOBW
 E  BG8.VOBW
 E  BG2.WTBW
 E  BG4.WTBW
 E  BG8.WTBW
 E  COF.A.BW
 E  CTF.A.BW
 E  CTP.A.BW
 E  CTF.ASBW
 E  CTF.C.BW
 E  CTF.E.BW
 E  CBF.F.BW
 E  CTF.FHBW
 E  CTF.FMBW
 E  CTF.FSBW
 E  CTF.FTBW
 E  CTP.FTBW
 E  CTF.F2BW
 E  CTF.H.BW
 E  CTF.I.BW
 E  CTF.IHBW
 E  CTF.ISBW
 E  CTF.LSBW
 E  CTF.PHBW
 E  CTF.PSBW
 E  CBF.P1BW
 E  CTC.P1BW
 E  CTF.P1BW
 E  CTV.P1BW
 E  CBF.P2BW
 E  CTC.P2BW
 E  CTF.P2BW
 E  CTV.P2BW
 E  CTC.P3BW
 E  CTF.P3BW
 E  CTV.P3BW
 E  CBF.2XBW
 E  CTC.2XBW
 E  CTV.2XBW
 G  GOI...BW
 G  NOI.G0BW
 G  NOI.G1BW
 G  NGS.G2BW
 G  NMO.G2BW
 G  NOI.G2BW
 G  NRO.G2BW
 G  NWU.G2BW
 G  N02.G2BW
 G  NOI.G3BW
 G  NGS.G4BW
 G  NMO.G4BW
 G  NOI.G4BW
 G  NRO.G4BW
 G  NWU.G4BW
 G  N02.G4BW
 G  NOI.G5BW
 G  NOI.G6BW
 G  NOI.G7BW
 G  NGS.G8BW
 G  NMO.G8BW
 G  NOI.G8BW
 G  NRO.G8BW
 G  NWU.G8BW
 G  N02.G8BW
 G  NOI.G9BW
 E  U.S...BW
 E  UFU...BW
 E  UH2...BW
 E  UNF...BW
 E  URX...BW
 E  U9R...BW
 E  WMO.G2BW
 E  WRO.G2BW
 E  WSU.G2BW
 E  WVP.G2BW
 E  W02.G2BW
 E  W12.G2BW
 E  W58.G2BW
 E  WMO.G4BW
 E  WRO.G4BW
 E  WSU.G4BW
 E  WVP.G4BW
 E  W02.G4BW
 E  W12.G4BW
 E  W58.G4BW
 E  WMO.G8BW
 E  WRO.G8BW
 E  WSU.G8BW
 E  WVP.G8BW
 E  W02.G8BW
 E  W12.G8BW
 E  W58.G8BW
 L  XCV.G2BW
 L  XSP.G2BW
 L  XVP.G2BW
 L  X12.G2BW
 L  X58.G2BW
 L  XCV.G4BW
 L  XSP.G4BW
 L  XVP.G4BW
 L  X12.G4BW
 L  X58.G4BW
 L  XCV.G8BW
 L  XSP.G8BW
 L  XVP.G8BW
 L  X12.G8BW
 L  X58.G8BW
COLUMNS
    BA3.G2BW  BA3...BW           -1.   BG2.VOBW            1.
    BA3.G2BW  BG2.WTBW          .685   NGS.G2BW          .685
    BA3.G2BW  WRO.G2BW      91.85397   WMO.G2BW      87.83771
    BA3.G2BW  W12.G2BW          35.2   W02.G2BW          94.4
    BA3.G2BW  WVP.G2BW            6.   W58.G2BW           6.2
    BA4.G2BW  BA4...BW           -1.   BG2.VOBW            1.
    BA4.G2BW  BG2.WTBW          .703   NGS.G2BW          .703
    BA4.G2BW  WRO.G2BW       95.0259   WMO.G2BW       90.4183
    BA4.G2BW  W12.G2BW          30.2   W02.G2BW           94.
    BA4.G2BW  WVP.G2BW            6.   W58.G2BW           -2.
    BA5.G2BW  BA5...BW           -1.   BG2.VOBW            1.
    BA5.G2BW  BG2.WTBW           .71   NGS.G2BW           .71
    BA5.G2BW  WRO.G2BW      93.37346   WMO.G2BW      89.22463
    BA5.G2BW  W12.G2BW           45.   W02.G2BW           96.
    BA5.G2BW  WVP.G2BW            6.   W58.G2BW          21.1
    BFD.G2BW  BFD...BW           -1.   BG2.VOBW            1.
    BFD.G2BW  BG2.WTBW          .786   NGS.G2BW          .786
    BFD.G2BW  WRO.G2BW      87.33868   WMO.G2BW      78.67732
    BFD.G2BW  W12.G2BW            2.   W02.G2BW           50.
    BFD.G2BW  WSU.G2BW        .09511   WVP.G2BW           2.1
    BFD.G2BW  W58.G2BW           -8.
    BFE.G2BW  BFE...BW           -1.   BG2.VOBW            1.
    BFE.G2BW  BG2.WTBW          .692   NGS.G2BW          .692
    BFE.G2BW  WRO.G2BW      92.03197   WMO.G2BW      81.05526
    BFE.G2BW  W12.G2BW           84.   W02.G2BW          100.
    BFE.G2BW  WSU.G2BW        .03204   WVP.G2BW           10.
    BFE.G2BW  W58.G2BW           47.
    BFF.G2BW  BFF...BW           -1.   BG2.VOBW            1.
    BFF.G2BW  BG2.WTBW           .78   NGS.G2BW           .78
    BFF.G2BW  WRO.G2BW      77.31816   WMO.G2BW      73.78616
    BFF.G2BW  W12.G2BW           10.   W02.G2BW           79.
    BFF.G2BW  WVP.G2BW            4.   W58.G2BW           -2.
    BFH.G2BW  BFH...BW           -1.   BG2.VOBW            1.
    BFH.G2BW  BG2.WTBW          .795   NGS.G2BW          .795
    BFH.G2BW  WRO.G2BW      87.42369   WMO.G2BW      78.28059
    BFH.G2BW  W12.G2BW            1.   W02.G2BW           44.
    BFH.G2BW  WSU.G2BW        .10891   WVP.G2BW            2.
    BFH.G2BW  W58.G2BW           -9.
    BFI.G2BW  BFI...BW           -1.   BG2.VOBW            1.
    BFI.G2BW  BG2.WTBW          .786   NGS.G2BW          .786
    BFI.G2BW  WRO.G2BW      81.33055   WMO.G2BW       76.0993
    BFI.G2BW  W12.G2BW            1.   W02.G2BW           44.
    BFI.G2BW  WSU.G2BW        .00534   WVP.G2BW           2.1
    BFI.G2BW  W58.G2BW           -9.
    BFJ.G2BW  BFJ...BW           -1.   BG2.VOBW            1.
    BFJ.G2BW  BG2.WTBW          .786   NGS.G2BW          .786
    BFJ.G2BW  WRO.G2BW      81.04273   WMO.G2BW      76.21092
    BFJ.G2BW  W12.G2BW            2.   W02.G2BW           50.
    BFJ.G2BW  WSU.G2BW        .00479   WVP.G2BW           2.1
    BFJ.G2BW  W58.G2BW           -8.
    BFK.G2BW  BFK...BW           -1.   BG2.VOBW            1.
    BFK.G2BW  BG2.WTBW          .805   NGS.G2BW          .805
    BFK.G2BW  WRO.G2BW      83.40638   WMO.G2BW      76.79616
    BFK.G2BW  W12.G2BW          -30.   W02.G2BW           20.
    BFK.G2BW  WSU.G2BW        .00064   WVP.G2BW            .5
    BFK.G2BW  W58.G2BW          -21.
    BFL.G2BW  BFL...BW           -1.   BG2.VOBW            1.
    BFL.G2BW  BG2.WTBW          .692   NGS.G2BW          .692
    BFL.G2BW  WRO.G2BW      92.39784   WMO.G2BW       81.5719
    BFL.G2BW  W12.G2BW           84.   W02.G2BW          100.
    BFL.G2BW  WSU.G2BW        .03204   WVP.G2BW           10.
    BFL.G2BW  W58.G2BW           47.
    BFM.G2BW  BFM...BW           -1.   BG2.VOBW            1.
    BFM.G2BW  BG2.WTBW          .692   NGS.G2BW          .692
    BFM.G2BW  WRO.G2BW      92.39784   WMO.G2BW       81.5719
    BFM.G2BW  W12.G2BW           84.   W02.G2BW          100.
    BFM.G2BW  WSU.G2BW        .01605   WVP.G2BW           10.
    BFM.G2BW  W58.G2BW           47.
    BFN.G2BW  BFN...BW           -1.   BG2.VOBW            1.
    BFN.G2BW  BG2.WTBW          .687   NGS.G2BW          .687
    BFN.G2BW  WRO.G2BW      92.70599   WMO.G2BW      81.63251
    BFN.G2BW  W12.G2BW           88.   W02.G2BW          100.
    BFN.G2BW  WSU.G2BW        .00756   WVP.G2BW          12.5
    BFN.G2BW  W58.G2BW           50.
    BFO.G2BW  BFO...BW           -1.   BG2.VOBW            1.
    BFO.G2BW  BG2.WTBW          .795   NGS.G2BW          .795
    BFO.G2BW  WRO.G2BW      85.02961   WMO.G2BW       75.6974
    BFO.G2BW  W12.G2BW            1.   W02.G2BW           44.
    BFO.G2BW  WSU.G2BW        .10891   WVP.G2BW            2.
    BFO.G2BW  W58.G2BW           -9.
    BFP.G2BW  BFP...BW           -1.   BG2.VOBW            1.
    BFP.G2BW  BG2.WTBW          .811   NGS.G2BW          .811
    BFP.G2BW  WRO.G2BW     105.15159   WMO.G2BW      90.22689
    BFP.G2BW  W12.G2BW           2.8   W02.G2BW          79.5
    BFP.G2BW  WVP.G2BW          2.43   W58.G2BW          -3.9
    BFQ.G2BW  BFQ...BW           -1.   BG2.VOBW            1.
    BFQ.G2BW  BG2.WTBW          .811   NGS.G2BW          .811
    BFQ.G2BW  WRO.G2BW     105.15159   WMO.G2BW      90.23941
    BFQ.G2BW  W12.G2BW           2.9   W02.G2BW           80.
    BFQ.G2BW  WVP.G2BW          2.43   W58.G2BW          -3.8
    BFS.G2BW  BFS...BW           -1.   BG2.VOBW            1.
    BFS.G2BW  BG2.WTBW          .687   NGS.G2BW          .687
    BFS.G2BW  WRO.G2BW      92.70599   WMO.G2BW      81.63251
    BFS.G2BW  W12.G2BW           88.   W02.G2BW          100.
    BFS.G2BW  WSU.G2BW        .01889   WVP.G2BW          12.5
    BFS.G2BW  W58.G2BW           50.
    BFT.G2BW  BFT...BW           -1.   BG2.VOBW            1.
    BFT.G2BW  BG2.WTBW           .78   NGS.G2BW           .78
    BFT.G2BW  WRO.G2BW      86.45545   WMO.G2BW      77.40834
    BFT.G2BW  W12.G2BW           10.   W02.G2BW           79.
    BFT.G2BW  WSU.G2BW         .0897   WVP.G2BW            4.
    BFT.G2BW  W58.G2BW           -2.
    BFU.G2BW  BFU...BW           -1.   BG2.VOBW            1.
    BFU.G2BW  BG2.WTBW          .805   NGS.G2BW          .805
    BFU.G2BW  WRO.G2BW      86.74963   WMO.G2BW      77.20759
    BFU.G2BW  W12.G2BW          -30.   W02.G2BW           20.
    BFU.G2BW  WSU.G2BW        .12477   WVP.G2BW            .5
    BFU.G2BW  W58.G2BW          -21.
    BFX.G2BW  BFX...BW           -1.   BG2.VOBW            1.
    BFX.G2BW  BG2.WTBW          .874   NGS.G2BW          .874
    BFX.G2BW  WRO.G2BW     102.56785   WMO.G2BW      87.99456
    BFX.G2BW  W12.G2BW          -31.   W02.G2BW           17.
    BFX.G2BW  WSU.G2BW         .0007   WVP.G2BW            .2
    BFX.G2BW  W58.G2BW          -21.
    BF1.G2BW  BF1...BW           -1.   BG2.VOBW            1.
    BF1.G2BW  BG2.WTBW          .835   NGS.G2BW          .835
    BF1.G2BW  WRO.G2BW     109.60858   WMO.G2BW       92.2351
    BF1.G2BW  W12.G2BW         -14.3   W02.G2BW          74.6
    BF1.G2BW  WVP.G2BW          1.02   W58.G2BW         -13.9
    BF2.G2BW  BF2...BW           -1.   BG2.VOBW            1.
    BF2.G2BW  BG2.WTBW          .736   NGS.G2BW          .736
    BF2.G2BW  WRO.G2BW      82.92617   WMO.G2BW      77.34037
    BF2.G2BW  W12.G2BW          79.6   W02.G2BW          100.
    BF2.G2BW  WVP.G2BW          2.43   W58.G2BW          -1.9
    BF3.G2BW  BF3...BW           -1.   BG2.VOBW            1.
    BF3.G2BW  BG2.WTBW          .835   NGS.G2BW          .835
    BF3.G2BW  WRO.G2BW     109.60858   WMO.G2BW      92.24962
    BF3.G2BW  W12.G2BW         -14.3   W02.G2BW          75.1
    BF3.G2BW  WVP.G2BW          1.02   W58.G2BW         -13.9
    BHC.G2BW  BHC...BW           -1.   BG2.VOBW            1.
    BHC.G2BW  BG2.WTBW          .772   NGS.G2BW          .772
    BHC.G2BW  WRO.G2BW     101.46649   WMO.G2BW      92.87805
    BHC.G2BW  W12.G2BW           22.   W02.G2BW           94.
    BHC.G2BW  WVP.G2BW           3.1   W58.G2BW            1.
    BIS.G2BW  BIS...BW           -1.   BG2.VOBW            1.
    BIS.G2BW  BG2.WTBW          .655   NGS.G2BW          .655
    BIS.G2BW  WRO.G2BW      90.09694   WMO.G2BW      86.13696
    BIS.G2BW  W12.G2BW          108.   W02.G2BW          100.
    BIS.G2BW  WVP.G2BW           13.   W58.G2BW           80.
    BIT.G2BW  BIT...BW           -1.   BG2.VOBW            1.
    BIT.G2BW  BG2.WTBW          .655   NGS.G2BW          .655
    BIT.G2BW  WRO.G2BW      88.88307   WMO.G2BW      84.80651
    BIT.G2BW  W12.G2BW          108.   W02.G2BW          100.
    BIT.G2BW  WVP.G2BW          13.2   W58.G2BW           80.
    BI4.G2BW  BI4...BW           -1.   BG2.VOBW            1.
    BI4.G2BW  BG2.WTBW          .563   NGS.G2BW          .563
    BI4.G2BW  WRO.G2BW     100.92828   WMO.G2BW      94.85789
    BI4.G2BW  W12.G2BW          120.   W02.G2BW          100.
    BI4.G2BW  WSU.G2BW        .00113   WVP.G2BW           78.
    BI4.G2BW  W58.G2BW          134.
    BI5.G2BW  BI5...BW           -1.   BG2.VOBW            1.
    BI5.G2BW  BG2.WTBW          .625   NGS.G2BW          .625
    BI5.G2BW  WRO.G2BW      92.91707   WMO.G2BW      87.95423
    BI5.G2BW  W12.G2BW          115.   W02.G2BW          100.
    BI5.G2BW  WSU.G2BW        .00156   WVP.G2BW           21.
    BI5.G2BW  W58.G2BW           95.
    BLF.G2BW  BLF...BW           -1.   BG2.VOBW            1.
    BLF.G2BW  BG2.WTBW          .658   NGS.G2BW          .658
    BLF.G2BW  WRO.G2BW      75.20602   WMO.G2BW      73.72027
    BLF.G2BW  W12.G2BW          110.   W02.G2BW          101.
    BLF.G2BW  WVP.G2BW          14.8   W58.G2BW           73.
    BLH.G2BW  BLH...BW           -1.   BG2.VOBW            1.
    BLH.G2BW  BG2.WTBW           .71   NGS.G2BW           .71
    BLH.G2BW  WRO.G2BW      70.06795   WMO.G2BW      69.42845
    BLH.G2BW  W12.G2BW           78.   W02.G2BW          100.
    BLH.G2BW  WSU.G2BW         .0213   WVP.G2BW           4.2
    BLH.G2BW  W58.G2BW           25.
    BLI.G2BW  BLI...BW           -1.   BG2.VOBW            1.
    BLI.G2BW  BG2.WTBW          .658   NGS.G2BW          .658
    BLI.G2BW  WRO.G2BW      79.88909   WMO.G2BW      78.12136
    BLI.G2BW  W12.G2BW          110.   W02.G2BW          101.
    BLI.G2BW  WSU.G2BW        .00132   WVP.G2BW          14.5
    BLI.G2BW  W58.G2BW           73.
    BLL.G2BW  BLL...BW           -1.   BG2.VOBW            1.
    BLL.G2BW  BG2.WTBW          .658   NGS.G2BW          .658
    BLL.G2BW  WRO.G2BW      75.20602   WMO.G2BW      73.72027
    BLL.G2BW  W12.G2BW          103.   W02.G2BW          101.
    BLL.G2BW  WSU.G2BW        .01448   WVP.G2BW          14.8
    BLL.G2BW  W58.G2BW           73.
    BLV.G2BW  BLV...BW           -1.   BG2.VOBW            1.
    BLV.G2BW  BG2.WTBW           .67   NGS.G2BW           .67
    BLV.G2BW  WRO.G2BW      74.42509   WMO.G2BW      73.02371
    BLV.G2BW  W12.G2BW           99.   W02.G2BW          103.
    BLV.G2BW  WSU.G2BW        .01675   WVP.G2BW          12.2
    BLV.G2BW  W58.G2BW           59.
    BN4.G2BW  BN4...BW           -1.   BG2.VOBW            1.
    BN4.G2BW  BG2.WTBW          .584   NGS.G2BW          .584
    BN4.G2BW  WRO.G2BW      94.11062   WMO.G2BW      88.35436
    BN4.G2BW  W12.G2BW          119.   W02.G2BW          100.
    BN4.G2BW  WSU.G2BW        .00117   WVP.G2BW           57.
    BN4.G2BW  W58.G2BW          122.
    BN5.G2BW  BN5...BW           -1.   BG2.VOBW            1.
    BN5.G2BW  BG2.WTBW          .631   NGS.G2BW          .631
    BN5.G2BW  WRO.G2BW      71.85843   WMO.G2BW      71.60925
    BN5.G2BW  W12.G2BW          121.   W02.G2BW           99.
    BN5.G2BW  WVP.G2BW           15.   W58.G2BW           86.
    BPE.G2BW  BPE...BW           -1.   BG2.VOBW            1.
    BPE.G2BW  BG2.WTBW          .669   NGS.G2BW          .669
    BPE.G2BW  WRO.G2BW      92.77312   WMO.G2BW      82.21873
    BPE.G2BW  W12.G2BW           98.   W02.G2BW          100.
    BPE.G2BW  WSU.G2BW        .01485   WVP.G2BW          13.5
    BPE.G2BW  W58.G2BW           58.
    BPQ.G2BW  BPQ...BW           -1.   BG2.VOBW            1.
    BPQ.G2BW  BG2.WTBW           .66   NGS.G2BW           .66
    BPQ.G2BW  WRO.G2BW      79.45782   WMO.G2BW      74.74378
    BPQ.G2BW  W12.G2BW          109.   W02.G2BW          100.
    BPQ.G2BW  WVP.G2BW          14.1   W58.G2BW          81.9
    BPT.G2BW  BPT...BW           -1.   BG2.VOBW            1.
    BPT.G2BW  BG2.WTBW          .669   NGS.G2BW          .669
    BPT.G2BW  WRO.G2BW      93.00495   WMO.G2BW      82.29867
    BPT.G2BW  W12.G2BW           98.   W02.G2BW          100.
    BPT.G2BW  WSU.G2BW        .01485   WVP.G2BW          13.5
    BPT.G2BW  W58.G2BW           58.
    BPU.G2BW  BPU...BW           -1.   BG2.VOBW            1.
    BPU.G2BW  BG2.WTBW          .693   NGS.G2BW          .693
    BPU.G2BW  WRO.G2BW      94.18285   WMO.G2BW      79.80062
    BPU.G2BW  W12.G2BW           93.   W02.G2BW          100.
    BPU.G2BW  WSU.G2BW         .0201   WVP.G2BW           13.
    BPU.G2BW  W58.G2BW           50.
    BPV.G2BW  BPV...BW           -1.   BG2.VOBW            1.
    BPV.G2BW  BG2.WTBW          .631   NGS.G2BW          .631
    BPV.G2BW  WRO.G2BW      76.03212   WMO.G2BW      73.15497
    BPV.G2BW  W12.G2BW           93.   W02.G2BW          100.
    BPV.G2BW  WSU.G2BW        .01956   WVP.G2BW           12.
    BPV.G2BW  W58.G2BW           48.
    BTO.G2BW  BTO...BW           -1.   BG2.VOBW            1.
    BTO.G2BW  BG2.WTBW          .872   NGS.G2BW          .872
    BTO.G2BW  WRO.G2BW       104.315   WMO.G2BW      91.00573
    BTO.G2BW  W12.G2BW           18.   W02.G2BW          100.
    BTO.G2BW  WSU.G2BW        .00044   WVP.G2BW           1.6
    BTO.G2BW  W58.G2BW          -20.
    BU4.G2BW  BU4...BW           -1.   BG2.VOBW            1.
    BU4.G2BW  BG2.WTBW           .62   NGS.G2BW           .62
    BU4.G2BW  WRO.G2BW     103.85185   WMO.G2BW      88.31541
    BU4.G2BW  W12.G2BW          113.   W02.G2BW          100.
    BU4.G2BW  WSU.G2BW        .00124   WVP.G2BW           61.
    BU4.G2BW  W58.G2BW         127.8
    BU5.G2BW  BU5...BW           -1.   BG2.VOBW            1.
    BU5.G2BW  BG2.WTBW          .659   NGS.G2BW          .659
    BU5.G2BW  WRO.G2BW     104.38066   WMO.G2BW       84.9222
    BU5.G2BW  W12.G2BW          115.   W02.G2BW          100.
    BU5.G2BW  WVP.G2BW          19.9   W58.G2BW          86.9
    BXA.G2BW  BXA...BW           -1.   BG2.VOBW            1.
    BXA.G2BW  BG2.WTBW          .796   NGS.G2BW          .796
    BXA.G2BW  WRO.G2BW      86.52499   WMO.G2BW      77.67402
    BXA.G2BW  W12.G2BW          -27.   W02.G2BW           28.
    BXA.G2BW  WSU.G2BW        .11303   WVP.G2BW            .7
    BXA.G2BW  W58.G2BW          -19.
    BXB.G2BW  BXB...BW           -1.   BG2.VOBW            1.
    BXB.G2BW  BG2.WTBW          .796   NGS.G2BW          .796
    BXB.G2BW  WRO.G2BW      83.15426   WMO.G2BW      77.10346
    BXB.G2BW  W12.G2BW          -27.   W02.G2BW           28.
    BXB.G2BW  WSU.G2BW        .00056   WVP.G2BW            .7
    BXB.G2BW  W58.G2BW          -19.
    BXC.G2BW  BXC...BW           -1.   BG2.VOBW            1.
    BXC.G2BW  BG2.WTBW          .839   NGS.G2BW          .839
    BXC.G2BW  WRO.G2BW     109.08673   WMO.G2BW      91.58015
    BXC.G2BW  W12.G2BW         -13.2   W02.G2BW          68.1
    BXC.G2BW  WVP.G2BW          1.42   W58.G2BW          -14.
    BXD.G2BW  BXD...BW           -1.   BG2.VOBW            1.
    BXD.G2BW  BG2.WTBW          .786   NGS.G2BW          .786
    BXD.G2BW  WRO.G2BW      85.53494   WMO.G2BW      76.62526
    BXD.G2BW  W12.G2BW            2.   W02.G2BW           50.
    BXD.G2BW  WSU.G2BW        .09511   WVP.G2BW           2.1
    BXD.G2BW  W58.G2BW           -8.
    BXE.G2BW  BXE...BW           -1.   BG2.VOBW            1.
    BXE.G2BW  BG2.WTBW          .723   NGS.G2BW          .723
    BXE.G2BW  WRO.G2BW      86.32131   WMO.G2BW      79.26925
    BXE.G2BW  W12.G2BW           79.   W02.G2BW          100.
    BXE.G2BW  WVP.G2BW          2.63   W58.G2BW           -.9
    BXF.G2BW  BXF...BW           -1.   BG2.VOBW            1.
    BXF.G2BW  BG2.WTBW           .78   NGS.G2BW           .78
    BXF.G2BW  WRO.G2BW      75.96553   WMO.G2BW      72.70699
    BXF.G2BW  W12.G2BW           10.   W02.G2BW           79.
    BXF.G2BW  WVP.G2BW            4.   W58.G2BW           -2.
    BXG.G2BW  BXG...BW           -1.   BG2.VOBW            1.
    BXG.G2BW  BG2.WTBW          .839   NGS.G2BW          .839
    BXG.G2BW  WRO.G2BW     109.08673   WMO.G2BW        91.595
    BXG.G2BW  W12.G2BW         -13.1   W02.G2BW          68.1
    BXG.G2BW  WVP.G2BW          1.42   W58.G2BW          -14.
    BXH.G2BW  BXH...BW           -1.   BG2.VOBW            1.
    BXH.G2BW  BG2.WTBW          .796   NGS.G2BW          .796
    BXH.G2BW  WRO.G2BW      83.73597   WMO.G2BW      74.14184
    BXH.G2BW  W12.G2BW          -27.   W02.G2BW           28.
    BXH.G2BW  WSU.G2BW        .11303   WVP.G2BW            .7
    BXH.G2BW  W58.G2BW          -19.
    BXI.G2BW  BXI...BW           -1.   BG2.VOBW            1.
    BXI.G2BW  BG2.WTBW          .795   NGS.G2BW          .795
    BXI.G2BW  WRO.G2BW      79.08957   WMO.G2BW      73.85299
    BXI.G2BW  W12.G2BW            1.   W02.G2BW           44.
    BXI.G2BW  WSU.G2BW        .00541   WVP.G2BW            2.
    BXI.G2BW  W58.G2BW           -9.
    BXJ.G2BW  BXJ...BW           -1.   BG2.VOBW            1.
    BXJ.G2BW  BG2.WTBW          .786   NGS.G2BW          .786
    BXJ.G2BW  WRO.G2BW      79.34372   WMO.G2BW      74.44831
    BXJ.G2BW  W12.G2BW            2.   W02.G2BW           50.
    BXJ.G2BW  WSU.G2BW        .00479   WVP.G2BW           2.1
    BXJ.G2BW  W58.G2BW           -8.
    BXK.G2BW  BXK...BW           -1.   BG2.VOBW            1.
    BXK.G2BW  BG2.WTBW          .805   NGS.G2BW          .805
    BXK.G2BW  WRO.G2BW      80.33046   WMO.G2BW      73.22415
    BXK.G2BW  W12.G2BW          -30.   W02.G2BW           20.
    BXK.G2BW  WSU.G2BW        .00064   WVP.G2BW            .5
    BXK.G2BW  W58.G2BW          -21.
    BXL.G2BW  BXL...BW           -1.   BG2.VOBW            1.
    BXL.G2BW  BG2.WTBW          .796   NGS.G2BW          .796
    BXL.G2BW  WRO.G2BW      80.50406   WMO.G2BW      73.95196
    BXL.G2BW  W12.G2BW          -27.   W02.G2BW           28.
    BXL.G2BW  WSU.G2BW        .00056   WVP.G2BW            .7
    BXL.G2BW  W58.G2BW          -19.
    BXM.G2BW  BXM...BW           -1.   BG2.VOBW            1.
    BXM.G2BW  BG2.WTBW          .692   NGS.G2BW          .692
    BXM.G2BW  WRO.G2BW      92.03197   WMO.G2BW      81.05526
    BXM.G2BW  W12.G2BW           84.   W02.G2BW          100.
    BXM.G2BW  WSU.G2BW        .01605   WVP.G2BW           10.
    BXM.G2BW  W58.G2BW           47.
    BXN.G2BW  BXN...BW           -1.   BG2.VOBW            1.
    BXN.G2BW  BG2.WTBW          .687   NGS.G2BW          .687
    BXN.G2BW  WRO.G2BW      92.09434   WMO.G2BW      81.45002
    BXN.G2BW  W12.G2BW           88.   W02.G2BW          100.
    BXN.G2BW  WSU.G2BW        .00756   WVP.G2BW          12.5
    BXN.G2BW  W58.G2BW           50.
    BXO.G2BW  BXO...BW           -1.   BG2.VOBW            1.
    BXO.G2BW  BG2.WTBW          .823   NGS.G2BW          .823
    BXO.G2BW  WRO.G2BW     102.34862   WMO.G2BW      88.75606
    BXO.G2BW  W12.G2BW           5.3   W02.G2BW           79.
    BXO.G2BW  WVP.G2BW          2.61   W58.G2BW          -2.8
    BXP.G2BW  BXP...BW           -1.   BG2.VOBW            1.
    BXP.G2BW  BG2.WTBW           .81   NGS.G2BW           .81
    BXP.G2BW  WRO.G2BW     105.15159   WMO.G2BW      90.21435
    BXP.G2BW  W12.G2BW           3.7   W02.G2BW          72.4
    BXP.G2BW  WVP.G2BW          2.71   W58.G2BW          -2.2
    BXQ.G2BW  BXQ...BW           -1.   BG2.VOBW            1.
    BXQ.G2BW  BG2.WTBW           .81   NGS.G2BW           .81
    BXQ.G2BW  WRO.G2BW     105.48297   WMO.G2BW      90.48006
    BXQ.G2BW  W12.G2BW           3.9   W02.G2BW          77.9
    BXQ.G2BW  WVP.G2BW          2.71   W58.G2BW          -2.1
    BXR.G2BW  BXR...BW           -1.   BG2.VOBW            1.
    BXR.G2BW  BG2.WTBW           .81   NGS.G2BW           .81
    BXR.G2BW  WRO.G2BW     105.81583   WMO.G2BW      90.58361
    BXR.G2BW  W12.G2BW            3.   W02.G2BW          77.2
    BXR.G2BW  WVP.G2BW           2.7   W58.G2BW          -2.3
    BXS.G2BW  BXS...BW           -1.   BG2.VOBW            1.
    BXS.G2BW  BG2.WTBW          .687   NGS.G2BW          .687
    BXS.G2BW  WRO.G2BW      92.09434   WMO.G2BW      81.45002
    BXS.G2BW  W12.G2BW           88.   W02.G2BW          100.
    BXS.G2BW  WSU.G2BW        .01889   WVP.G2BW          12.5
    BXS.G2BW  W58.G2BW           50.
    BXT.G2BW  BXT...BW           -1.   BG2.VOBW            1.
    BXT.G2BW  BG2.WTBW           .78   NGS.G2BW           .78
    BXT.G2BW  WRO.G2BW      84.71223   WMO.G2BW       76.0144
    BXT.G2BW  W12.G2BW           10.   W02.G2BW           79.
    BXT.G2BW  WSU.G2BW         .0897   WVP.G2BW            4.
    BXT.G2BW  W58.G2BW           -2.
    BXU.G2BW  BXU...BW           -1.   BG2.VOBW            1.
    BXU.G2BW  BG2.WTBW          .805   NGS.G2BW          .805
    BXU.G2BW  WRO.G2BW      83.48678   WMO.G2BW      73.24984
    BXU.G2BW  W12.G2BW          -30.   W02.G2BW           20.
    BXU.G2BW  WSU.G2BW        .12477   WVP.G2BW            .5
    BXU.G2BW  W58.G2BW          -21.
    BXX.G2BW  BXX...BW           -1.   BG2.VOBW            1.
    BXX.G2BW  BG2.WTBW          .874   NGS.G2BW          .874
    BXX.G2BW  WRO.G2BW     102.56785   WMO.G2BW      87.99456
    BXX.G2BW  W12.G2BW          -31.   W02.G2BW           17.
    BXX.G2BW  WSU.G2BW         .0007   WVP.G2BW            .2
    BXX.G2BW  W58.G2BW          -21.
    BX1.G2BW  BX1...BW           -1.   BG2.VOBW            1.
    BX1.G2BW  BG2.WTBW          .753   NGS.G2BW          .753
    BX1.G2BW  WRO.G2BW      74.14662   WMO.G2BW      71.09379
    BX1.G2BW  W12.G2BW          86.2   W02.G2BW          100.
    BX1.G2BW  WVP.G2BW          2.35   W58.G2BW          -1.9
    BX2.G2BW  BX2...BW           -1.   BG2.VOBW            1.
    BX2.G2BW  BG2.WTBW          .717   NGS.G2BW          .717
    BX2.G2BW  WRO.G2BW       89.0308   WMO.G2BW      80.78009
    BX2.G2BW  W12.G2BW          76.4   W02.G2BW          100.
    BX2.G2BW  WVP.G2BW          2.68   W58.G2BW           -.8
    BX3.G2BW  BX3...BW           -1.   BG2.VOBW            1.
    BX3.G2BW  BG2.WTBW          .852   NGS.G2BW          .852
    BX3.G2BW  WRO.G2BW     105.65125   WMO.G2BW      90.05458
    BX3.G2BW  W12.G2BW         -12.6   W02.G2BW           73.
    BX3.G2BW  WVP.G2BW          1.37   W58.G2BW          -14.
    BX4.G2BW  BX4...BW           -1.   BG2.VOBW            1.
    BX4.G2BW  BG2.WTBW          .838   NGS.G2BW          .838
    BX4.G2BW  WRO.G2BW     109.64842   WMO.G2BW      91.83832
    BX4.G2BW  W12.G2BW         -13.4   W02.G2BW          71.3
    BX4.G2BW  WVP.G2BW          1.48   W58.G2BW          -14.
    BX6.G2BW  BX6...BW           -1.   BG2.VOBW            1.
    BX6.G2BW  BG2.WTBW          .693   NGS.G2BW          .693
    BX6.G2BW  WRO.G2BW      93.74377   WMO.G2BW      79.60009
    BX6.G2BW  W12.G2BW           93.   W02.G2BW          100.
    BX6.G2BW  WSU.G2BW         .0201   WVP.G2BW           13.
    BX6.G2BW  W58.G2BW           50.
    B1A.G2BW  B1A...BW           -1.   BG2.VOBW            1.
    B1A.G2BW  BG2.WTBW          .693   NGS.G2BW          .693
    B1A.G2BW  WRO.G2BW      94.90093   WMO.G2BW      84.96442
    B1A.G2BW  W12.G2BW           71.   W02.G2BW          100.
    B1A.G2BW  WVP.G2BW          5.56   W58.G2BW          23.5
    B1B.G2BW  B1B...BW           -1.   BG2.VOBW            1.
    B1B.G2BW  BG2.WTBW          .806   NGS.G2BW          .806
    B1B.G2BW  WRO.G2BW      99.30716   WMO.G2BW      87.42257
    B1B.G2BW  W12.G2BW         20.77   W02.G2BW           82.
    B1B.G2BW  WVP.G2BW          3.47   W58.G2BW          3.55
    B1C.G2BW  B1C...BW           -1.   BG2.VOBW            1.
    B1C.G2BW  BG2.WTBW          .845   NGS.G2BW          .845
    B1C.G2BW  WRO.G2BW      98.90249   WMO.G2BW      86.84432
    B1C.G2BW  W12.G2BW          2.67   W02.G2BW          22.1
    B1C.G2BW  WVP.G2BW          2.98   W58.G2BW          -2.4
    B1D.G2BW  B1D...BW           -1.   BG2.VOBW            1.
    B1D.G2BW  BG2.WTBW          .823   NGS.G2BW          .823
    B1D.G2BW  WRO.G2BW     103.69118   WMO.G2BW      89.42014
    B1D.G2BW  W12.G2BW           2.9   W02.G2BW          79.8
    B1D.G2BW  WVP.G2BW          2.36   W58.G2BW          -4.4
    B1G.G2BW  B1G...BW           -1.   BG2.VOBW            1.
    B1G.G2BW  BG2.WTBW          .688   NGS.G2BW          .688
    B1G.G2BW  WRO.G2BW      100.1228   WMO.G2BW      88.21309
    B1G.G2BW  W12.G2BW          64.3   W02.G2BW          100.
    B1G.G2BW  WVP.G2BW          5.52   W58.G2BW          22.2
    B1H.G2BW  B1H...BW           -1.   BG2.VOBW            1.
    B1H.G2BW  BG2.WTBW            .8   NGS.G2BW            .8
    B1H.G2BW  WRO.G2BW     102.96806   WMO.G2BW      89.46263
    B1H.G2BW  W12.G2BW          16.5   W02.G2BW          79.8
    B1H.G2BW  WVP.G2BW          3.61   W58.G2BW           4.4
    B1I.G2BW  B1I...BW           -1.   BG2.VOBW            1.
    B1I.G2BW  BG2.WTBW          .841   NGS.G2BW          .841
    B1I.G2BW  WRO.G2BW     101.95318   WMO.G2BW      88.53996
    B1I.G2BW  W12.G2BW           2.9   W02.G2BW          19.9
    B1I.G2BW  WVP.G2BW          3.09   W58.G2BW          -1.6
    B1J.G2BW  B1J...BW           -1.   BG2.VOBW            1.
    B1J.G2BW  BG2.WTBW           .81   NGS.G2BW           .81
    B1J.G2BW  WRO.G2BW     106.96742   WMO.G2BW      91.25379
    B1J.G2BW  W12.G2BW            .7   W02.G2BW          78.8
    B1J.G2BW  WVP.G2BW          2.42   W58.G2BW          -4.1
    B1K.G2BW  B1K...BW           -1.   BG2.VOBW            1.
    B1K.G2BW  BG2.WTBW          .693   NGS.G2BW          .693
    B1K.G2BW  WRO.G2BW      94.90093   WMO.G2BW      84.96442
    B1K.G2BW  W12.G2BW           71.   W02.G2BW          100.
    B1K.G2BW  WVP.G2BW          5.56   W58.G2BW          23.5
    B1L.G2BW  B1L...BW           -1.   BG2.VOBW            1.
    B1L.G2BW  BG2.WTBW          .806   NGS.G2BW          .806
    B1L.G2BW  WRO.G2BW      99.30716   WMO.G2BW      87.42257
    B1L.G2BW  W12.G2BW         20.77   W02.G2BW           82.
    B1L.G2BW  WVP.G2BW          3.47   W58.G2BW          3.55
    B1M.G2BW  B1M...BW           -1.   BG2.VOBW            1.
    B1M.G2BW  BG2.WTBW          .845   NGS.G2BW          .845
    B1M.G2BW  WRO.G2BW      98.90249   WMO.G2BW      86.84432
    B1M.G2BW  W12.G2BW          2.67   W02.G2BW          22.1
    B1M.G2BW  WVP.G2BW          2.98   W58.G2BW          -2.4
    B1N.G2BW  B1N...BW           -1.   BG2.VOBW            1.
    B1N.G2BW  BG2.WTBW          .823   NGS.G2BW          .823
    B1N.G2BW  WRO.G2BW     103.69118   WMO.G2BW      89.42014
    B1N.G2BW  W12.G2BW           2.9   W02.G2BW          79.8
    B1N.G2BW  WVP.G2BW          2.36   W58.G2BW          -4.4
    B1O.G2BW  B1O...BW           -1.   BG2.VOBW            1.
    B1O.G2BW  BG2.WTBW          .823   NGS.G2BW          .823
    B1O.G2BW  WRO.G2BW     102.34862   WMO.G2BW      88.75606
    B1O.G2BW  W12.G2BW           5.3   W02.G2BW           79.
    B1O.G2BW  WVP.G2BW          2.61   W58.G2BW          -2.8
    B2A.G2BW  B2A...BW           -1.   BG2.VOBW            1.
    B2A.G2BW  BG2.WTBW           .74   NGS.G2BW           .74
    B2A.G2BW  WRO.G2BW     108.32812   WMO.G2BW      92.94314
    B2A.G2BW  W12.G2BW          19.7   W02.G2BW          95.2
    B2A.G2BW  WVP.G2BW          1.18   W58.G2BW          -9.4
    B2B.G2BW  B2B...BW           -1.   BG2.VOBW            1.
    B2B.G2BW  BG2.WTBW          .848   NGS.G2BW          .848
    B2B.G2BW  WRO.G2BW     102.76529   WMO.G2BW      88.35055
    B2B.G2BW  W12.G2BW           -7.   W02.G2BW          73.8
    B2B.G2BW  WVP.G2BW          1.17   W58.G2BW         -13.4
    B2C.G2BW  B2C...BW           -1.   BG2.VOBW            1.
    B2C.G2BW  BG2.WTBW          .885   NGS.G2BW          .885
    B2C.G2BW  WRO.G2BW     100.04377   WMO.G2BW      86.52755
    B2C.G2BW  W12.G2BW         -16.7   W02.G2BW           7.3
    B2C.G2BW  WVP.G2BW          1.38   W58.G2BW         -14.7
    B2D.G2BW  B2D...BW           -1.   BG2.VOBW            1.
    B2D.G2BW  BG2.WTBW          .847   NGS.G2BW          .847
    B2D.G2BW  WRO.G2BW     107.43624   WMO.G2BW       91.2576
    B2D.G2BW  W12.G2BW          -15.   W02.G2BW          75.1
    B2D.G2BW  WVP.G2BW          1.02   W58.G2BW         -13.9
    B2G.G2BW  B2G...BW           -1.   BG2.VOBW            1.
    B2G.G2BW  BG2.WTBW          .738   NGS.G2BW          .738
    B2G.G2BW  WRO.G2BW     114.92604   WMO.G2BW      97.63681
    B2G.G2BW  W12.G2BW          14.1   W02.G2BW          100.
    B2G.G2BW  WVP.G2BW          1.13   W58.G2BW          -9.7
    B2H.G2BW  B2H...BW           -1.   BG2.VOBW            1.
    B2H.G2BW  BG2.WTBW          .844   NGS.G2BW          .844
    B2H.G2BW  WRO.G2BW       108.951   WMO.G2BW       92.2581
    B2H.G2BW  W12.G2BW          -9.2   W02.G2BW          71.1
    B2H.G2BW  WVP.G2BW          1.32   W58.G2BW         -13.5
    B2I.G2BW  B2I...BW           -1.   BG2.VOBW            1.
    B2I.G2BW  BG2.WTBW          .883   NGS.G2BW          .883
    B2I.G2BW  WRO.G2BW     103.65702   WMO.G2BW      88.58483
    B2I.G2BW  W12.G2BW         -16.9   W02.G2BW           7.5
    B2I.G2BW  WVP.G2BW          1.35   W58.G2BW         -14.6
    B2J.G2BW  B2J...BW           -1.   BG2.VOBW            1.
    B2J.G2BW  BG2.WTBW          .834   NGS.G2BW          .834
    B2J.G2BW  WRO.G2BW     110.38305   WMO.G2BW      92.88876
    B2J.G2BW  W12.G2BW         -15.9   W02.G2BW          74.3
    B2J.G2BW  WVP.G2BW          1.01   W58.G2BW         -13.9
    B2K.G2BW  B2K...BW           -1.   BG2.VOBW            1.
    B2K.G2BW  BG2.WTBW           .74   NGS.G2BW           .74
    B2K.G2BW  WRO.G2BW     108.32812   WMO.G2BW      92.94314
    B2K.G2BW  W12.G2BW          19.7   W02.G2BW          95.2
    B2K.G2BW  WVP.G2BW          1.18   W58.G2BW          -9.4
    B2L.G2BW  B2L...BW           -1.   BG2.VOBW            1.
    B2L.G2BW  BG2.WTBW          .848   NGS.G2BW          .848
    B2L.G2BW  WRO.G2BW     102.76529   WMO.G2BW      88.35055
    B2L.G2BW  W12.G2BW           -7.   W02.G2BW          73.8
    B2L.G2BW  WVP.G2BW          1.17   W58.G2BW         -13.4
    B2M.G2BW  B2M...BW           -1.   BG2.VOBW            1.
    B2M.G2BW  BG2.WTBW          .885   NGS.G2BW          .885
    B2M.G2BW  WRO.G2BW     100.04377   WMO.G2BW      86.52755
    B2M.G2BW  W12.G2BW         -16.7   W02.G2BW           7.3
    B2M.G2BW  WVP.G2BW          1.38   W58.G2BW         -14.7
    B2N.G2BW  B2N...BW           -1.   BG2.VOBW            1.
    B2N.G2BW  BG2.WTBW          .847   NGS.G2BW          .847
    B2N.G2BW  WRO.G2BW     107.43624   WMO.G2BW       91.2576
    B2N.G2BW  W12.G2BW          -15.   W02.G2BW          75.1
    B2N.G2BW  WVP.G2BW          1.02   W58.G2BW         -13.9
    B2O.G2BW  B2O...BW           -1.   BG2.VOBW            1.
    B2O.G2BW  BG2.WTBW          .852   NGS.G2BW          .852
    B2O.G2BW  WRO.G2BW     105.65125   WMO.G2BW      90.05458
    B2O.G2BW  W12.G2BW         -12.6   W02.G2BW           73.
    B2O.G2BW  WVP.G2BW          1.37   W58.G2BW          -14.
    B4U.G2BW  B4U...BW           -1.   BG2.VOBW            1.
    B4U.G2BW  BG2.WTBW           .62   NGS.G2BW           .62
    B4U.G2BW  WRO.G2BW     103.85185   WMO.G2BW      88.31541
    B4U.G2BW  W12.G2BW          113.   W02.G2BW          100.
    B4U.G2BW  WSU.G2BW        .00124   WVP.G2BW           61.
    B4U.G2BW  W58.G2BW         127.8
    B5A.G2BW  B5A...BW           -1.   BG2.VOBW            1.
    B5A.G2BW  BG2.WTBW          .684   NGS.G2BW          .684
    B5A.G2BW  WRO.G2BW      92.06722   WMO.G2BW      82.21907
    B5A.G2BW  W12.G2BW          93.8   W02.G2BW          100.
    B5A.G2BW  WVP.G2BW          3.47   W58.G2BW           1.2
    B5B.G2BW  B5B...BW           -1.   BG2.VOBW            1.
    B5B.G2BW  BG2.WTBW          .756   NGS.G2BW          .756
    B5B.G2BW  WRO.G2BW      73.17229   WMO.G2BW      71.98974
    B5B.G2BW  W12.G2BW          92.1   W02.G2BW          100.
    B5B.G2BW  WVP.G2BW          2.36   W58.G2BW         -2.33
    B5C.G2BW  B5C...BW           -1.   BG2.VOBW            1.
    B5C.G2BW  BG2.WTBW          .752   NGS.G2BW          .752
    B5C.G2BW  WRO.G2BW      71.31302   WMO.G2BW      71.11826
    B5C.G2BW  W12.G2BW          99.2   W02.G2BW          100.
    B5C.G2BW  WVP.G2BW          2.71   W58.G2BW           .82
    B5D.G2BW  B5D...BW           -1.   BG2.VOBW            1.
    B5D.G2BW  BG2.WTBW           .75   NGS.G2BW           .75
    B5D.G2BW  WRO.G2BW      76.53985   WMO.G2BW      71.71782
    B5D.G2BW  W12.G2BW          82.2   W02.G2BW          100.
    B5D.G2BW  WVP.G2BW          2.28   W58.G2BW          -2.3
    B5G.G2BW  B5G...BW           -1.   BG2.VOBW            1.
    B5G.G2BW  BG2.WTBW          .666   NGS.G2BW          .666
    B5G.G2BW  WRO.G2BW     102.10471   WMO.G2BW         87.97
    B5G.G2BW  W12.G2BW           91.   W02.G2BW          100.
    B5G.G2BW  WVP.G2BW          3.65   W58.G2BW            .9
    B5H.G2BW  B5H...BW           -1.   BG2.VOBW            1.
    B5H.G2BW  BG2.WTBW          .735   NGS.G2BW          .735
    B5H.G2BW  WRO.G2BW      80.53522   WMO.G2BW      75.95929
    B5H.G2BW  W12.G2BW          85.7   W02.G2BW          100.
    B5H.G2BW  WVP.G2BW          2.55   W58.G2BW          -1.4
    B5I.G2BW  B5I...BW           -1.   BG2.VOBW            1.
    B5I.G2BW  BG2.WTBW          .743   NGS.G2BW          .743
    B5I.G2BW  WRO.G2BW      73.10682   WMO.G2BW      71.95043
    B5I.G2BW  W12.G2BW           98.   W02.G2BW          100.
    B5I.G2BW  WVP.G2BW          2.83   W58.G2BW           1.5
    B5J.G2BW  B5J...BW           -1.   BG2.VOBW            1.
    B5J.G2BW  BG2.WTBW          .715   NGS.G2BW          .715
    B5J.G2BW  WRO.G2BW      92.52513   WMO.G2BW      82.83746
    B5J.G2BW  W12.G2BW          70.2   W02.G2BW          100.
    B5J.G2BW  WVP.G2BW          2.58   W58.G2BW          -1.6
    B5K.G2BW  B5K...BW           -1.   BG2.VOBW            1.
    B5K.G2BW  BG2.WTBW          .684   NGS.G2BW          .684
    B5K.G2BW  WRO.G2BW      92.06722   WMO.G2BW      82.21907
    B5K.G2BW  W12.G2BW          93.8   W02.G2BW          100.
    B5K.G2BW  WVP.G2BW          3.47   W58.G2BW           1.2
    B5L.G2BW  B5L...BW           -1.   BG2.VOBW            1.
    B5L.G2BW  BG2.WTBW          .756   NGS.G2BW          .756
    B5L.G2BW  WRO.G2BW      73.17229   WMO.G2BW      71.98974
    B5L.G2BW  W12.G2BW          92.1   W02.G2BW          100.
    B5L.G2BW  WVP.G2BW          2.36   W58.G2BW         -2.33
    B5M.G2BW  B5M...BW           -1.   BG2.VOBW            1.
    B5M.G2BW  BG2.WTBW          .752   NGS.G2BW          .752
    B5M.G2BW  WRO.G2BW      71.31302   WMO.G2BW      71.11826
    B5M.G2BW  W12.G2BW          99.2   W02.G2BW          100.
    B5M.G2BW  WVP.G2BW          2.71   W58.G2BW           .82
    B5N.G2BW  B5N...BW           -1.   BG2.VOBW            1.
    B5N.G2BW  BG2.WTBW           .75   NGS.G2BW           .75
    B5N.G2BW  WRO.G2BW      76.53985   WMO.G2BW      71.71782
    B5N.G2BW  W12.G2BW          82.2   W02.G2BW          100.
    B5N.G2BW  WVP.G2BW          2.28   W58.G2BW          -2.3
    B5O.G2BW  B5O...BW           -1.   BG2.VOBW            1.
    B5O.G2BW  BG2.WTBW          .753   NGS.G2BW          .753
    B5O.G2BW  WRO.G2BW      74.14662   WMO.G2BW      71.09379
    B5O.G2BW  W12.G2BW          86.2   W02.G2BW          100.
    B5O.G2BW  WVP.G2BW          2.35   W58.G2BW          -1.9
    B5S.G2BW  B5S...BW           -1.   BG2.VOBW            1.
    B5S.G2BW  BG2.WTBW          .625   NGS.G2BW          .625
    B5S.G2BW  WRO.G2BW      93.25986   WMO.G2BW      88.88656
    B5S.G2BW  W12.G2BW          122.   W02.G2BW          100.
    B5S.G2BW  WSU.G2BW        .00156   WVP.G2BW           21.
    B5S.G2BW  W58.G2BW           96.
    BA3.G4BW  BA3...BW           -1.   BG4.VOBW            1.
    BA3.G4BW  BG4.WTBW          .685   NGS.G4BW          .685
    BA3.G4BW  WRO.G4BW      91.85397   WMO.G4BW      87.83771
    BA3.G4BW  W12.G4BW          35.2   W02.G4BW          94.4
    BA3.G4BW  WVP.G4BW            6.   W58.G4BW           6.2
    BA4.G4BW  BA4...BW           -1.   BG4.VOBW            1.
    BA4.G4BW  BG4.WTBW          .703   NGS.G4BW          .703
    BA4.G4BW  WRO.G4BW       95.0259   WMO.G4BW       90.4183
    BA4.G4BW  W12.G4BW          30.2   W02.G4BW           94.
    BA4.G4BW  WVP.G4BW            6.   W58.G4BW           -2.
    BA5.G4BW  BA5...BW           -1.   BG4.VOBW            1.
    BA5.G4BW  BG4.WTBW           .71   NGS.G4BW           .71
    BA5.G4BW  WRO.G4BW      93.37346   WMO.G4BW      89.22463
    BA5.G4BW  W12.G4BW           45.   W02.G4BW           96.
    BA5.G4BW  WVP.G4BW            6.   W58.G4BW          21.1
    BFD.G4BW  BFD...BW           -1.   BG4.VOBW            1.
    BFD.G4BW  BG4.WTBW          .786   NGS.G4BW          .786
    BFD.G4BW  WRO.G4BW      87.33868   WMO.G4BW      78.67732
    BFD.G4BW  W12.G4BW            2.   W02.G4BW           50.
    BFD.G4BW  WSU.G4BW        .09511   WVP.G4BW           2.1
    BFD.G4BW  W58.G4BW           -8.
    BFE.G4BW  BFE...BW           -1.   BG4.VOBW            1.
    BFE.G4BW  BG4.WTBW          .692   NGS.G4BW          .692
    BFE.G4BW  WRO.G4BW      92.03197   WMO.G4BW      81.05526
    BFE.G4BW  W12.G4BW           84.   W02.G4BW          100.
    BFE.G4BW  WSU.G4BW        .03204   WVP.G4BW           10.
    BFE.G4BW  W58.G4BW           47.
    BFF.G4BW  BFF...BW           -1.   BG4.VOBW            1.
    BFF.G4BW  BG4.WTBW           .78   NGS.G4BW           .78
    BFF.G4BW  WRO.G4BW      77.31816   WMO.G4BW      73.78616
    BFF.G4BW  W12.G4BW           10.   W02.G4BW           79.
    BFF.G4BW  WVP.G4BW            4.   W58.G4BW           -2.
    BFH.G4BW  BFH...BW           -1.   BG4.VOBW            1.
    BFH.G4BW  BG4.WTBW          .795   NGS.G4BW          .795
    BFH.G4BW  WRO.G4BW      87.42369   WMO.G4BW      78.28059
    BFH.G4BW  W12.G4BW            1.   W02.G4BW           44.
    BFH.G4BW  WSU.G4BW        .10891   WVP.G4BW            2.
    BFH.G4BW  W58.G4BW           -9.
    BFI.G4BW  BFI...BW           -1.   BG4.VOBW            1.
    BFI.G4BW  BG4.WTBW          .786   NGS.G4BW          .786
    BFI.G4BW  WRO.G4BW      81.33055   WMO.G4BW       76.0993
    BFI.G4BW  W12.G4BW            1.   W02.G4BW           44.
    BFI.G4BW  WSU.G4BW        .00534   WVP.G4BW           2.1
    BFI.G4BW  W58.G4BW           -9.
    BFJ.G4BW  BFJ...BW           -1.   BG4.VOBW            1.
    BFJ.G4BW  BG4.WTBW          .786   NGS.G4BW          .786
    BFJ.G4BW  WRO.G4BW      81.04273   WMO.G4BW      76.21092
    BFJ.G4BW  W12.G4BW            2.   W02.G4BW           50.
    BFJ.G4BW  WSU.G4BW        .00479   WVP.G4BW           2.1
    BFJ.G4BW  W58.G4BW           -8.
    BFK.G4BW  BFK...BW           -1.   BG4.VOBW            1.
    BFK.G4BW  BG4.WTBW          .805   NGS.G4BW          .805
    BFK.G4BW  WRO.G4BW      83.40638   WMO.G4BW      76.79616
    BFK.G4BW  W12.G4BW          -30.   W02.G4BW           20.
    BFK.G4BW  WSU.G4BW        .00064   WVP.G4BW            .5
    BFK.G4BW  W58.G4BW          -21.
    BFL.G4BW  BFL...BW           -1.   BG4.VOBW            1.
    BFL.G4BW  BG4.WTBW          .692   NGS.G4BW          .692
    BFL.G4BW  WRO.G4BW      92.39784   WMO.G4BW       81.5719
    BFL.G4BW  W12.G4BW           84.   W02.G4BW          100.
    BFL.G4BW  WSU.G4BW        .03204   WVP.G4BW           10.
    BFL.G4BW  W58.G4BW           47.
    BFM.G4BW  BFM...BW           -1.   BG4.VOBW            1.
    BFM.G4BW  BG4.WTBW          .692   NGS.G4BW          .692
    BFM.G4BW  WRO.G4BW      92.39784   WMO.G4BW       81.5719
    BFM.G4BW  W12.G4BW           84.   W02.G4BW          100.
    BFM.G4BW  WSU.G4BW        .01605   WVP.G4BW           10.
    BFM.G4BW  W58.G4BW           47.
    BFN.G4BW  BFN...BW           -1.   BG4.VOBW            1.
    BFN.G4BW  BG4.WTBW          .687   NGS.G4BW          .687
    BFN.G4BW  WRO.G4BW      92.70599   WMO.G4BW      81.63251
    BFN.G4BW  W12.G4BW           88.   W02.G4BW          100.
    BFN.G4BW  WSU.G4BW        .00756   WVP.G4BW          12.5
    BFN.G4BW  W58.G4BW           50.
    BFO.G4BW  BFO...BW           -1.   BG4.VOBW            1.
    BFO.G4BW  BG4.WTBW          .795   NGS.G4BW          .795
    BFO.G4BW  WRO.G4BW      85.02961   WMO.G4BW       75.6974
    BFO.G4BW  W12.G4BW            1.   W02.G4BW           44.
    BFO.G4BW  WSU.G4BW        .10891   WVP.G4BW            2.
    BFO.G4BW  W58.G4BW           -9.
    BFP.G4BW  BFP...BW           -1.   BG4.VOBW            1.
    BFP.G4BW  BG4.WTBW          .811   NGS.G4BW          .811
    BFP.G4BW  WRO.G4BW     105.15159   WMO.G4BW      90.22689
    BFP.G4BW  W12.G4BW           2.8   W02.G4BW          79.5
    BFP.G4BW  WVP.G4BW          2.43   W58.G4BW          -3.9
    BFQ.G4BW  BFQ...BW           -1.   BG4.VOBW            1.
    BFQ.G4BW  BG4.WTBW          .811   NGS.G4BW          .811
    BFQ.G4BW  WRO.G4BW     105.15159   WMO.G4BW      90.23941
    BFQ.G4BW  W12.G4BW           2.9   W02.G4BW           80.
    BFQ.G4BW  WVP.G4BW          2.43   W58.G4BW          -3.8
    BFS.G4BW  BFS...BW           -1.   BG4.VOBW            1.
    BFS.G4BW  BG4.WTBW          .687   NGS.G4BW          .687
    BFS.G4BW  WRO.G4BW      92.70599   WMO.G4BW      81.63251
    BFS.G4BW  W12.G4BW           88.   W02.G4BW          100.
    BFS.G4BW  WSU.G4BW        .01889   WVP.G4BW          12.5
    BFS.G4BW  W58.G4BW           50.
    BFT.G4BW  BFT...BW           -1.   BG4.VOBW            1.
    BFT.G4BW  BG4.WTBW           .78   NGS.G4BW           .78
    BFT.G4BW  WRO.G4BW      86.45545   WMO.G4BW      77.40834
    BFT.G4BW  W12.G4BW           10.   W02.G4BW           79.
    BFT.G4BW  WSU.G4BW         .0897   WVP.G4BW            4.
    BFT.G4BW  W58.G4BW           -2.
    BFU.G4BW  BFU...BW           -1.   BG4.VOBW            1.
    BFU.G4BW  BG4.WTBW          .805   NGS.G4BW          .805
    BFU.G4BW  WRO.G4BW      86.74963   WMO.G4BW      77.20759
    BFU.G4BW  W12.G4BW          -30.   W02.G4BW           20.
    BFU.G4BW  WSU.G4BW        .12477   WVP.G4BW            .5
    BFU.G4BW  W58.G4BW          -21.
    BFX.G4BW  BFX...BW           -1.   BG4.VOBW            1.
    BFX.G4BW  BG4.WTBW          .874   NGS.G4BW          .874
    BFX.G4BW  WRO.G4BW     102.56785   WMO.G4BW      87.99456
    BFX.G4BW  W12.G4BW          -31.   W02.G4BW           17.
    BFX.G4BW  WSU.G4BW         .0007   WVP.G4BW            .2
    BFX.G4BW  W58.G4BW          -21.
    BF1.G4BW  BF1...BW           -1.   BG4.VOBW            1.
    BF1.G4BW  BG4.WTBW          .835   NGS.G4BW          .835
    BF1.G4BW  WRO.G4BW     109.60858   WMO.G4BW       92.2351
    BF1.G4BW  W12.G4BW         -14.3   W02.G4BW          74.6
    BF1.G4BW  WVP.G4BW          1.02   W58.G4BW         -13.9
    BF2.G4BW  BF2...BW           -1.   BG4.VOBW            1.
    BF2.G4BW  BG4.WTBW          .736   NGS.G4BW          .736
    BF2.G4BW  WRO.G4BW      82.92617   WMO.G4BW      77.34037
    BF2.G4BW  W12.G4BW          79.6   W02.G4BW          100.
    BF2.G4BW  WVP.G4BW          2.43   W58.G4BW          -1.9
    BF3.G4BW  BF3...BW           -1.   BG4.VOBW            1.
    BF3.G4BW  BG4.WTBW          .835   NGS.G4BW          .835
    BF3.G4BW  WRO.G4BW     109.60858   WMO.G4BW      92.24962
    BF3.G4BW  W12.G4BW         -14.3   W02.G4BW          75.1
    BF3.G4BW  WVP.G4BW          1.02   W58.G4BW         -13.9
    BHC.G4BW  BHC...BW           -1.   BG4.VOBW            1.
    BHC.G4BW  BG4.WTBW          .772   NGS.G4BW          .772
    BHC.G4BW  WRO.G4BW     101.46649   WMO.G4BW      92.87805
    BHC.G4BW  W12.G4BW           22.   W02.G4BW           94.
    BHC.G4BW  WVP.G4BW           3.1   W58.G4BW            1.
    BIS.G4BW  BIS...BW           -1.   BG4.VOBW            1.
    BIS.G4BW  BG4.WTBW          .655   NGS.G4BW          .655
    BIS.G4BW  WRO.G4BW      90.09694   WMO.G4BW      86.13696
    BIS.G4BW  W12.G4BW          108.   W02.G4BW          100.
    BIS.G4BW  WVP.G4BW           13.   W58.G4BW           80.
    BIT.G4BW  BIT...BW           -1.   BG4.VOBW            1.
    BIT.G4BW  BG4.WTBW          .655   NGS.G4BW          .655
    BIT.G4BW  WRO.G4BW      88.88307   WMO.G4BW      84.80651
    BIT.G4BW  W12.G4BW          108.   W02.G4BW          100.
    BIT.G4BW  WVP.G4BW          13.2   W58.G4BW           80.
    BI4.G4BW  BI4...BW           -1.   BG4.VOBW            1.
    BI4.G4BW  BG4.WTBW          .563   NGS.G4BW          .563
    BI4.G4BW  WRO.G4BW     100.92828   WMO.G4BW      94.85789
    BI4.G4BW  W12.G4BW          120.   W02.G4BW          100.
    BI4.G4BW  WSU.G4BW        .00113   WVP.G4BW           78.
    BI4.G4BW  W58.G4BW          134.
    BI5.G4BW  BI5...BW           -1.   BG4.VOBW            1.
    BI5.G4BW  BG4.WTBW          .625   NGS.G4BW          .625
    BI5.G4BW  WRO.G4BW      92.91707   WMO.G4BW      87.95423
    BI5.G4BW  W12.G4BW          115.   W02.G4BW          100.
    BI5.G4BW  WSU.G4BW        .00156   WVP.G4BW           21.
    BI5.G4BW  W58.G4BW           95.
    BLF.G4BW  BLF...BW           -1.   BG4.VOBW            1.
    BLF.G4BW  BG4.WTBW          .658   NGS.G4BW          .658
    BLF.G4BW  WRO.G4BW      75.20602   WMO.G4BW      73.72027
    BLF.G4BW  W12.G4BW          110.   W02.G4BW          101.
    BLF.G4BW  WVP.G4BW          14.8   W58.G4BW           73.
    BLH.G4BW  BLH...BW           -1.   BG4.VOBW            1.
    BLH.G4BW  BG4.WTBW           .71   NGS.G4BW           .71
    BLH.G4BW  WRO.G4BW      70.06795   WMO.G4BW      69.42845
    BLH.G4BW  W12.G4BW           78.   W02.G4BW          100.
    BLH.G4BW  WSU.G4BW         .0213   WVP.G4BW           4.2
    BLH.G4BW  W58.G4BW           25.
    BLI.G4BW  BLI...BW           -1.   BG4.VOBW            1.
    BLI.G4BW  BG4.WTBW          .658   NGS.G4BW          .658
    BLI.G4BW  WRO.G4BW      79.88909   WMO.G4BW      78.12136
    BLI.G4BW  W12.G4BW          110.   W02.G4BW          101.
    BLI.G4BW  WSU.G4BW        .00132   WVP.G4BW          14.5
    BLI.G4BW  W58.G4BW           73.
    BLL.G4BW  BLL...BW           -1.   BG4.VOBW            1.
    BLL.G4BW  BG4.WTBW          .658   NGS.G4BW          .658
    BLL.G4BW  WRO.G4BW      75.20602   WMO.G4BW      73.72027
    BLL.G4BW  W12.G4BW          103.   W02.G4BW          101.
    BLL.G4BW  WSU.G4BW        .01448   WVP.G4BW          14.8
    BLL.G4BW  W58.G4BW           73.
    BLV.G4BW  BLV...BW           -1.   BG4.VOBW            1.
    BLV.G4BW  BG4.WTBW           .67   NGS.G4BW           .67
    BLV.G4BW  WRO.G4BW      74.42509   WMO.G4BW      73.02371
    BLV.G4BW  W12.G4BW           99.   W02.G4BW          103.
    BLV.G4BW  WSU.G4BW        .01675   WVP.G4BW          12.2
    BLV.G4BW  W58.G4BW           59.
    BN4.G4BW  BN4...BW           -1.   BG4.VOBW            1.
    BN4.G4BW  BG4.WTBW          .584   NGS.G4BW          .584
    BN4.G4BW  WRO.G4BW      94.11062   WMO.G4BW      88.35436
    BN4.G4BW  W12.G4BW          119.   W02.G4BW          100.
    BN4.G4BW  WSU.G4BW        .00117   WVP.G4BW           57.
    BN4.G4BW  W58.G4BW          122.
    BN5.G4BW  BN5...BW           -1.   BG4.VOBW            1.
    BN5.G4BW  BG4.WTBW          .631   NGS.G4BW          .631
    BN5.G4BW  WRO.G4BW      71.85843   WMO.G4BW      71.60925
    BN5.G4BW  W12.G4BW          121.   W02.G4BW           99.
    BN5.G4BW  WVP.G4BW           15.   W58.G4BW           86.
    BPE.G4BW  BPE...BW           -1.   BG4.VOBW            1.
    BPE.G4BW  BG4.WTBW          .669   NGS.G4BW          .669
    BPE.G4BW  WRO.G4BW      92.77312   WMO.G4BW      82.21873
    BPE.G4BW  W12.G4BW           98.   W02.G4BW          100.
    BPE.G4BW  WSU.G4BW        .01485   WVP.G4BW          13.5
    BPE.G4BW  W58.G4BW           58.
    BPQ.G4BW  BPQ...BW           -1.   BG4.VOBW            1.
    BPQ.G4BW  BG4.WTBW           .66   NGS.G4BW           .66
    BPQ.G4BW  WRO.G4BW      79.45782   WMO.G4BW      74.74378
    BPQ.G4BW  W12.G4BW          109.   W02.G4BW          100.
    BPQ.G4BW  WVP.G4BW          14.1   W58.G4BW          81.9
    BPT.G4BW  BPT...BW           -1.   BG4.VOBW            1.
    BPT.G4BW  BG4.WTBW          .669   NGS.G4BW          .669
    BPT.G4BW  WRO.G4BW      93.00495   WMO.G4BW      82.29867
    BPT.G4BW  W12.G4BW           98.   W02.G4BW          100.
    BPT.G4BW  WSU.G4BW        .01485   WVP.G4BW          13.5
    BPT.G4BW  W58.G4BW           58.
    BPU.G4BW  BPU...BW           -1.   BG4.VOBW            1.
    BPU.G4BW  BG4.WTBW          .693   NGS.G4BW          .693
    BPU.G4BW  WRO.G4BW      94.18285   WMO.G4BW      79.80062
    BPU.G4BW  W12.G4BW           93.   W02.G4BW          100.
    BPU.G4BW  WSU.G4BW         .0201   WVP.G4BW           13.
    BPU.G4BW  W58.G4BW           50.
    BPV.G4BW  BPV...BW           -1.   BG4.VOBW            1.
    BPV.G4BW  BG4.WTBW          .631   NGS.G4BW          .631
    BPV.G4BW  WRO.G4BW      76.03212   WMO.G4BW      73.15497
    BPV.G4BW  W12.G4BW           93.   W02.G4BW          100.
    BPV.G4BW  WSU.G4BW        .01956   WVP.G4BW           12.
    BPV.G4BW  W58.G4BW           48.
    BTO.G4BW  BTO...BW           -1.   BG4.VOBW            1.
    BTO.G4BW  BG4.WTBW          .872   NGS.G4BW          .872
    BTO.G4BW  WRO.G4BW       104.315   WMO.G4BW      91.00573
    BTO.G4BW  W12.G4BW           18.   W02.G4BW          100.
    BTO.G4BW  WSU.G4BW        .00044   WVP.G4BW           1.6
    BTO.G4BW  W58.G4BW          -20.
    BU4.G4BW  BU4...BW           -1.   BG4.VOBW            1.
    BU4.G4BW  BG4.WTBW           .62   NGS.G4BW           .62
    BU4.G4BW  WRO.G4BW     103.85185   WMO.G4BW      88.31541
    BU4.G4BW  W12.G4BW          113.   W02.G4BW          100.
    BU4.G4BW  WSU.G4BW        .00124   WVP.G4BW           61.
    BU4.G4BW  W58.G4BW         127.8
    BU5.G4BW  BU5...BW           -1.   BG4.VOBW            1.
    BU5.G4BW  BG4.WTBW          .659   NGS.G4BW          .659
    BU5.G4BW  WRO.G4BW     104.38066   WMO.G4BW       84.9222
    BU5.G4BW  W12.G4BW          115.   W02.G4BW          100.
    BU5.G4BW  WVP.G4BW          19.9   W58.G4BW          86.9
    BXA.G4BW  BXA...BW           -1.   BG4.VOBW            1.
    BXA.G4BW  BG4.WTBW          .796   NGS.G4BW          .796
    BXA.G4BW  WRO.G4BW      86.52499   WMO.G4BW      77.67402
    BXA.G4BW  W12.G4BW          -27.   W02.G4BW           28.
    BXA.G4BW  WSU.G4BW        .11303   WVP.G4BW            .7
    BXA.G4BW  W58.G4BW          -19.
    BXB.G4BW  BXB...BW           -1.   BG4.VOBW            1.
    BXB.G4BW  BG4.WTBW          .796   NGS.G4BW          .796
    BXB.G4BW  WRO.G4BW      83.15426   WMO.G4BW      77.10346
    BXB.G4BW  W12.G4BW          -27.   W02.G4BW           28.
    BXB.G4BW  WSU.G4BW        .00056   WVP.G4BW            .7
    BXB.G4BW  W58.G4BW          -19.
    BXC.G4BW  BXC...BW           -1.   BG4.VOBW            1.
    BXC.G4BW  BG4.WTBW          .839   NGS.G4BW          .839
    BXC.G4BW  WRO.G4BW     109.08673   WMO.G4BW      91.58015
    BXC.G4BW  W12.G4BW         -13.2   W02.G4BW          68.1
    BXC.G4BW  WVP.G4BW          1.42   W58.G4BW          -14.
    BXD.G4BW  BXD...BW           -1.   BG4.VOBW            1.
    BXD.G4BW  BG4.WTBW          .786   NGS.G4BW          .786
    BXD.G4BW  WRO.G4BW      85.53494   WMO.G4BW      76.62526
    BXD.G4BW  W12.G4BW            2.   W02.G4BW           50.
    BXD.G4BW  WSU.G4BW        .09511   WVP.G4BW           2.1
    BXD.G4BW  W58.G4BW           -8.
    BXE.G4BW  BXE...BW           -1.   BG4.VOBW            1.
    BXE.G4BW  BG4.WTBW          .723   NGS.G4BW          .723
    BXE.G4BW  WRO.G4BW      86.32131   WMO.G4BW      79.26925
    BXE.G4BW  W12.G4BW           79.   W02.G4BW          100.
    BXE.G4BW  WVP.G4BW          2.63   W58.G4BW           -.9
    BXF.G4BW  BXF...BW           -1.   BG4.VOBW            1.
    BXF.G4BW  BG4.WTBW           .78   NGS.G4BW           .78
    BXF.G4BW  WRO.G4BW      75.96553   WMO.G4BW      72.70699
    BXF.G4BW  W12.G4BW           10.   W02.G4BW           79.
    BXF.G4BW  WVP.G4BW            4.   W58.G4BW           -2.
    BXG.G4BW  BXG...BW           -1.   BG4.VOBW            1.
    BXG.G4BW  BG4.WTBW          .839   NGS.G4BW          .839
    BXG.G4BW  WRO.G4BW     109.08673   WMO.G4BW        91.595
    BXG.G4BW  W12.G4BW         -13.1   W02.G4BW          68.1
    BXG.G4BW  WVP.G4BW          1.42   W58.G4BW          -14.
    BXH.G4BW  BXH...BW           -1.   BG4.VOBW            1.
    BXH.G4BW  BG4.WTBW          .796   NGS.G4BW          .796
    BXH.G4BW  WRO.G4BW      83.73597   WMO.G4BW      74.14184
    BXH.G4BW  W12.G4BW          -27.   W02.G4BW           28.
    BXH.G4BW  WSU.G4BW        .11303   WVP.G4BW            .7
    BXH.G4BW  W58.G4BW          -19.
    BXI.G4BW  BXI...BW           -1.   BG4.VOBW            1.
    BXI.G4BW  BG4.WTBW          .795   NGS.G4BW          .795
    BXI.G4BW  WRO.G4BW      79.08957   WMO.G4BW      73.85299
    BXI.G4BW  W12.G4BW            1.   W02.G4BW           44.
    BXI.G4BW  WSU.G4BW        .00541   WVP.G4BW            2.
    BXI.G4BW  W58.G4BW           -9.
    BXJ.G4BW  BXJ...BW           -1.   BG4.VOBW            1.
    BXJ.G4BW  BG4.WTBW          .786   NGS.G4BW          .786
    BXJ.G4BW  WRO.G4BW      79.34372   WMO.G4BW      74.44831
    BXJ.G4BW  W12.G4BW            2.   W02.G4BW           50.
    BXJ.G4BW  WSU.G4BW        .00479   WVP.G4BW           2.1
    BXJ.G4BW  W58.G4BW           -8.
    BXK.G4BW  BXK...BW           -1.   BG4.VOBW            1.
    BXK.G4BW  BG4.WTBW          .805   NGS.G4BW          .805
    BXK.G4BW  WRO.G4BW      80.33046   WMO.G4BW      73.22415
    BXK.G4BW  W12.G4BW          -30.   W02.G4BW           20.
    BXK.G4BW  WSU.G4BW        .00064   WVP.G4BW            .5
    BXK.G4BW  W58.G4BW          -21.
    BXL.G4BW  BXL...BW           -1.   BG4.VOBW            1.
    BXL.G4BW  BG4.WTBW          .796   NGS.G4BW          .796
    BXL.G4BW  WRO.G4BW      80.50406   WMO.G4BW      73.95196
    BXL.G4BW  W12.G4BW          -27.   W02.G4BW           28.
    BXL.G4BW  WSU.G4BW        .00056   WVP.G4BW            .7
    BXL.G4BW  W58.G4BW          -19.
    BXM.G4BW  BXM...BW           -1.   BG4.VOBW            1.
    BXM.G4BW  BG4.WTBW          .692   NGS.G4BW          .692
    BXM.G4BW  WRO.G4BW      92.03197   WMO.G4BW      81.05526
    BXM.G4BW  W12.G4BW           84.   W02.G4BW          100.
    BXM.G4BW  WSU.G4BW        .01605   WVP.G4BW           10.
    BXM.G4BW  W58.G4BW           47.
    BXN.G4BW  BXN...BW           -1.   BG4.VOBW            1.
    BXN.G4BW  BG4.WTBW          .687   NGS.G4BW          .687
    BXN.G4BW  WRO.G4BW      92.09434   WMO.G4BW      81.45002
    BXN.G4BW  W12.G4BW           88.   W02.G4BW          100.
    BXN.G4BW  WSU.G4BW        .00756   WVP.G4BW          12.5
    BXN.G4BW  W58.G4BW           50.
    BXO.G4BW  BXO...BW           -1.   BG4.VOBW            1.
    BXO.G4BW  BG4.WTBW          .823   NGS.G4BW          .823
    BXO.G4BW  WRO.G4BW     102.34862   WMO.G4BW      88.75606
    BXO.G4BW  W12.G4BW           5.3   W02.G4BW           79.
    BXO.G4BW  WVP.G4BW          2.61   W58.G4BW          -2.8
    BXP.G4BW  BXP...BW           -1.   BG4.VOBW            1.
    BXP.G4BW  BG4.WTBW           .81   NGS.G4BW           .81
    BXP.G4BW  WRO.G4BW     105.15159   WMO.G4BW      90.21435
    BXP.G4BW  W12.G4BW           3.7   W02.G4BW          72.4
    BXP.G4BW  WVP.G4BW          2.71   W58.G4BW          -2.2
    BXQ.G4BW  BXQ...BW           -1.   BG4.VOBW            1.
    BXQ.G4BW  BG4.WTBW           .81   NGS.G4BW           .81
    BXQ.G4BW  WRO.G4BW     105.48297   WMO.G4BW      90.48006
    BXQ.G4BW  W12.G4BW           3.9   W02.G4BW          77.9
    BXQ.G4BW  WVP.G4BW          2.71   W58.G4BW          -2.1
    BXR.G4BW  BXR...BW           -1.   BG4.VOBW            1.
    BXR.G4BW  BG4.WTBW           .81   NGS.G4BW           .81
    BXR.G4BW  WRO.G4BW     105.81583   WMO.G4BW      90.58361
    BXR.G4BW  W12.G4BW            3.   W02.G4BW          77.2
    BXR.G4BW  WVP.G4BW           2.7   W58.G4BW          -2.3
    BXS.G4BW  BXS...BW           -1.   BG4.VOBW            1.
    BXS.G4BW  BG4.WTBW          .687   NGS.G4BW          .687
    BXS.G4BW  WRO.G4BW      92.09434   WMO.G4BW      81.45002
    BXS.G4BW  W12.G4BW           88.   W02.G4BW          100.
    BXS.G4BW  WSU.G4BW        .01889   WVP.G4BW          12.5
    BXS.G4BW  W58.G4BW           50.
    BXT.G4BW  BXT...BW           -1.   BG4.VOBW            1.
    BXT.G4BW  BG4.WTBW           .78   NGS.G4BW           .78
    BXT.G4BW  WRO.G4BW      84.71223   WMO.G4BW       76.0144
    BXT.G4BW  W12.G4BW           10.   W02.G4BW           79.
    BXT.G4BW  WSU.G4BW         .0897   WVP.G4BW            4.
    BXT.G4BW  W58.G4BW           -2.
    BXU.G4BW  BXU...BW           -1.   BG4.VOBW            1.
    BXU.G4BW  BG4.WTBW          .805   NGS.G4BW          .805
    BXU.G4BW  WRO.G4BW      83.48678   WMO.G4BW      73.24984
    BXU.G4BW  W12.G4BW          -30.   W02.G4BW           20.
    BXU.G4BW  WSU.G4BW        .12477   WVP.G4BW            .5
    BXU.G4BW  W58.G4BW          -21.
    BXX.G4BW  BXX...BW           -1.   BG4.VOBW            1.
    BXX.G4BW  BG4.WTBW          .874   NGS.G4BW          .874
    BXX.G4BW  WRO.G4BW     102.56785   WMO.G4BW      87.99456
    BXX.G4BW  W12.G4BW          -31.   W02.G4BW           17.
    BXX.G4BW  WSU.G4BW         .0007   WVP.G4BW            .2
    BXX.G4BW  W58.G4BW          -21.
    BX1.G4BW  BX1...BW           -1.   BG4.VOBW            1.
    BX1.G4BW  BG4.WTBW          .753   NGS.G4BW          .753
    BX1.G4BW  WRO.G4BW      74.14662   WMO.G4BW      71.09379
    BX1.G4BW  W12.G4BW          86.2   W02.G4BW          100.
    BX1.G4BW  WVP.G4BW          2.35   W58.G4BW          -1.9
    BX2.G4BW  BX2...BW           -1.   BG4.VOBW            1.
    BX2.G4BW  BG4.WTBW          .717   NGS.G4BW          .717
    BX2.G4BW  WRO.G4BW       89.0308   WMO.G4BW      80.78009
    BX2.G4BW  W12.G4BW          76.4   W02.G4BW          100.
    BX2.G4BW  WVP.G4BW          2.68   W58.G4BW           -.8
    BX3.G4BW  BX3...BW           -1.   BG4.VOBW            1.
    BX3.G4BW  BG4.WTBW          .852   NGS.G4BW          .852
    BX3.G4BW  WRO.G4BW     105.65125   WMO.G4BW      90.05458
    BX3.G4BW  W12.G4BW         -12.6   W02.G4BW           73.
    BX3.G4BW  WVP.G4BW          1.37   W58.G4BW          -14.
    BX4.G4BW  BX4...BW           -1.   BG4.VOBW            1.
    BX4.G4BW  BG4.WTBW          .838   NGS.G4BW          .838
    BX4.G4BW  WRO.G4BW     109.64842   WMO.G4BW      91.83832
    BX4.G4BW  W12.G4BW         -13.4   W02.G4BW          71.3
    BX4.G4BW  WVP.G4BW          1.48   W58.G4BW          -14.
    BX6.G4BW  BX6...BW           -1.   BG4.VOBW            1.
    BX6.G4BW  BG4.WTBW          .693   NGS.G4BW          .693
    BX6.G4BW  WRO.G4BW      93.74377   WMO.G4BW      79.60009
    BX6.G4BW  W12.G4BW           93.   W02.G4BW          100.
    BX6.G4BW  WSU.G4BW         .0201   WVP.G4BW           13.
    BX6.G4BW  W58.G4BW           50.
    B1A.G4BW  B1A...BW           -1.   BG4.VOBW            1.
    B1A.G4BW  BG4.WTBW          .693   NGS.G4BW          .693
    B1A.G4BW  WRO.G4BW      94.90093   WMO.G4BW      84.96442
    B1A.G4BW  W12.G4BW           71.   W02.G4BW          100.
    B1A.G4BW  WVP.G4BW          5.56   W58.G4BW          23.5
    B1B.G4BW  B1B...BW           -1.   BG4.VOBW            1.
    B1B.G4BW  BG4.WTBW          .806   NGS.G4BW          .806
    B1B.G4BW  WRO.G4BW      99.30716   WMO.G4BW      87.42257
    B1B.G4BW  W12.G4BW         20.77   W02.G4BW           82.
    B1B.G4BW  WVP.G4BW          3.47   W58.G4BW          3.55
    B1C.G4BW  B1C...BW           -1.   BG4.VOBW            1.
    B1C.G4BW  BG4.WTBW          .845   NGS.G4BW          .845
    B1C.G4BW  WRO.G4BW      98.90249   WMO.G4BW      86.84432
    B1C.G4BW  W12.G4BW          2.67   W02.G4BW          22.1
    B1C.G4BW  WVP.G4BW          2.98   W58.G4BW          -2.4
    B1D.G4BW  B1D...BW           -1.   BG4.VOBW            1.
    B1D.G4BW  BG4.WTBW          .823   NGS.G4BW          .823
    B1D.G4BW  WRO.G4BW     103.69118   WMO.G4BW      89.42014
    B1D.G4BW  W12.G4BW           2.9   W02.G4BW          79.8
    B1D.G4BW  WVP.G4BW          2.36   W58.G4BW          -4.4
    B1G.G4BW  B1G...BW           -1.   BG4.VOBW            1.
    B1G.G4BW  BG4.WTBW          .688   NGS.G4BW          .688
    B1G.G4BW  WRO.G4BW      100.1228   WMO.G4BW      88.21309
    B1G.G4BW  W12.G4BW          64.3   W02.G4BW          100.
    B1G.G4BW  WVP.G4BW          5.52   W58.G4BW          22.2
    B1H.G4BW  B1H...BW           -1.   BG4.VOBW            1.
    B1H.G4BW  BG4.WTBW            .8   NGS.G4BW            .8
    B1H.G4BW  WRO.G4BW     102.96806   WMO.G4BW      89.46263
    B1H.G4BW  W12.G4BW          16.5   W02.G4BW          79.8
    B1H.G4BW  WVP.G4BW          3.61   W58.G4BW           4.4
    B1I.G4BW  B1I...BW           -1.   BG4.VOBW            1.
    B1I.G4BW  BG4.WTBW          .841   NGS.G4BW          .841
    B1I.G4BW  WRO.G4BW     101.95318   WMO.G4BW      88.53996
    B1I.G4BW  W12.G4BW           2.9   W02.G4BW          19.9
    B1I.G4BW  WVP.G4BW          3.09   W58.G4BW          -1.6
    B1J.G4BW  B1J...BW           -1.   BG4.VOBW            1.
    B1J.G4BW  BG4.WTBW           .81   NGS.G4BW           .81
    B1J.G4BW  WRO.G4BW     106.96742   WMO.G4BW      91.25379
    B1J.G4BW  W12.G4BW            .7   W02.G4BW          78.8
    B1J.G4BW  WVP.G4BW          2.42   W58.G4BW          -4.1
    B1K.G4BW  B1K...BW           -1.   BG4.VOBW            1.
    B1K.G4BW  BG4.WTBW          .693   NGS.G4BW          .693
    B1K.G4BW  WRO.G4BW      94.90093   WMO.G4BW      84.96442
    B1K.G4BW  W12.G4BW           71.   W02.G4BW          100.
    B1K.G4BW  WVP.G4BW          5.56   W58.G4BW          23.5
    B1L.G4BW  B1L...BW           -1.   BG4.VOBW            1.
    B1L.G4BW  BG4.WTBW          .806   NGS.G4BW          .806
    B1L.G4BW  WRO.G4BW      99.30716   WMO.G4BW      87.42257
    B1L.G4BW  W12.G4BW         20.77   W02.G4BW           82.
    B1L.G4BW  WVP.G4BW          3.47   W58.G4BW          3.55
    B1M.G4BW  B1M...BW           -1.   BG4.VOBW            1.
    B1M.G4BW  BG4.WTBW          .845   NGS.G4BW          .845
    B1M.G4BW  WRO.G4BW      98.90249   WMO.G4BW      86.84432
    B1M.G4BW  W12.G4BW          2.67   W02.G4BW          22.1
    B1M.G4BW  WVP.G4BW          2.98   W58.G4BW          -2.4
    B1N.G4BW  B1N...BW           -1.   BG4.VOBW            1.
    B1N.G4BW  BG4.WTBW          .823   NGS.G4BW          .823
    B1N.G4BW  WRO.G4BW     103.69118   WMO.G4BW      89.42014
    B1N.G4BW  W12.G4BW           2.9   W02.G4BW          79.8
    B1N.G4BW  WVP.G4BW          2.36   W58.G4BW          -4.4
    B1O.G4BW  B1O...BW           -1.   BG4.VOBW            1.
    B1O.G4BW  BG4.WTBW          .823   NGS.G4BW          .823
    B1O.G4BW  WRO.G4BW     102.34862   WMO.G4BW      88.75606
    B1O.G4BW  W12.G4BW           5.3   W02.G4BW           79.
    B1O.G4BW  WVP.G4BW          2.61   W58.G4BW          -2.8
    B2A.G4BW  B2A...BW           -1.   BG4.VOBW            1.
    B2A.G4BW  BG4.WTBW           .74   NGS.G4BW           .74
    B2A.G4BW  WRO.G4BW     108.32812   WMO.G4BW      92.94314
    B2A.G4BW  W12.G4BW          19.7   W02.G4BW          95.2
    B2A.G4BW  WVP.G4BW          1.18   W58.G4BW          -9.4
    B2B.G4BW  B2B...BW           -1.   BG4.VOBW            1.
    B2B.G4BW  BG4.WTBW          .848   NGS.G4BW          .848
    B2B.G4BW  WRO.G4BW     102.76529   WMO.G4BW      88.35055
    B2B.G4BW  W12.G4BW           -7.   W02.G4BW          73.8
    B2B.G4BW  WVP.G4BW          1.17   W58.G4BW         -13.4
    B2C.G4BW  B2C...BW           -1.   BG4.VOBW            1.
    B2C.G4BW  BG4.WTBW          .885   NGS.G4BW          .885
    B2C.G4BW  WRO.G4BW     100.04377   WMO.G4BW      86.52755
    B2C.G4BW  W12.G4BW         -16.7   W02.G4BW           7.3
    B2C.G4BW  WVP.G4BW          1.38   W58.G4BW         -14.7
    B2D.G4BW  B2D...BW           -1.   BG4.VOBW            1.
    B2D.G4BW  BG4.WTBW          .847   NGS.G4BW          .847
    B2D.G4BW  WRO.G4BW     107.43624   WMO.G4BW       91.2576
    B2D.G4BW  W12.G4BW          -15.   W02.G4BW          75.1
    B2D.G4BW  WVP.G4BW          1.02   W58.G4BW         -13.9
    B2G.G4BW  B2G...BW           -1.   BG4.VOBW            1.
    B2G.G4BW  BG4.WTBW          .738   NGS.G4BW          .738
    B2G.G4BW  WRO.G4BW     114.92604   WMO.G4BW      97.63681
    B2G.G4BW  W12.G4BW          14.1   W02.G4BW          100.
    B2G.G4BW  WVP.G4BW          1.13   W58.G4BW          -9.7
    B2H.G4BW  B2H...BW           -1.   BG4.VOBW            1.
    B2H.G4BW  BG4.WTBW          .844   NGS.G4BW          .844
    B2H.G4BW  WRO.G4BW       108.951   WMO.G4BW       92.2581
    B2H.G4BW  W12.G4BW          -9.2   W02.G4BW          71.1
    B2H.G4BW  WVP.G4BW          1.32   W58.G4BW         -13.5
    B2I.G4BW  B2I...BW           -1.   BG4.VOBW            1.
    B2I.G4BW  BG4.WTBW          .883   NGS.G4BW          .883
    B2I.G4BW  WRO.G4BW     103.65702   WMO.G4BW      88.58483
    B2I.G4BW  W12.G4BW         -16.9   W02.G4BW           7.5
    B2I.G4BW  WVP.G4BW          1.35   W58.G4BW         -14.6
    B2J.G4BW  B2J...BW           -1.   BG4.VOBW            1.
    B2J.G4BW  BG4.WTBW          .834   NGS.G4BW          .834
    B2J.G4BW  WRO.G4BW     110.38305   WMO.G4BW      92.88876
    B2J.G4BW  W12.G4BW         -15.9   W02.G4BW          74.3
    B2J.G4BW  WVP.G4BW          1.01   W58.G4BW         -13.9
    B2K.G4BW  B2K...BW           -1.   BG4.VOBW            1.
    B2K.G4BW  BG4.WTBW           .74   NGS.G4BW           .74
    B2K.G4BW  WRO.G4BW     108.32812   WMO.G4BW      92.94314
    B2K.G4BW  W12.G4BW          19.7   W02.G4BW          95.2
    B2K.G4BW  WVP.G4BW          1.18   W58.G4BW          -9.4
    B2L.G4BW  B2L...BW           -1.   BG4.VOBW            1.
    B2L.G4BW  BG4.WTBW          .848   NGS.G4BW          .848
    B2L.G4BW  WRO.G4BW     102.76529   WMO.G4BW      88.35055
    B2L.G4BW  W12.G4BW           -7.   W02.G4BW          73.8
    B2L.G4BW  WVP.G4BW          1.17   W58.G4BW         -13.4
    B2M.G4BW  B2M...BW           -1.   BG4.VOBW            1.
    B2M.G4BW  BG4.WTBW          .885   NGS.G4BW          .885
    B2M.G4BW  WRO.G4BW     100.04377   WMO.G4BW      86.52755
    B2M.G4BW  W12.G4BW         -16.7   W02.G4BW           7.3
    B2M.G4BW  WVP.G4BW          1.38   W58.G4BW         -14.7
    B2N.G4BW  B2N...BW           -1.   BG4.VOBW            1.
    B2N.G4BW  BG4.WTBW          .847   NGS.G4BW          .847
    B2N.G4BW  WRO.G4BW     107.43624   WMO.G4BW       91.2576
    B2N.G4BW  W12.G4BW          -15.   W02.G4BW          75.1
    B2N.G4BW  WVP.G4BW          1.02   W58.G4BW         -13.9
    B2O.G4BW  B2O...BW           -1.   BG4.VOBW            1.
    B2O.G4BW  BG4.WTBW          .852   NGS.G4BW          .852
    B2O.G4BW  WRO.G4BW     105.65125   WMO.G4BW      90.05458
    B2O.G4BW  W12.G4BW         -12.6   W02.G4BW           73.
    B2O.G4BW  WVP.G4BW          1.37   W58.G4BW          -14.
    B4U.G4BW  B4U...BW           -1.   BG4.VOBW            1.
    B4U.G4BW  BG4.WTBW           .62   NGS.G4BW           .62
    B4U.G4BW  WRO.G4BW     103.85185   WMO.G4BW      88.31541
    B4U.G4BW  W12.G4BW          113.   W02.G4BW          100.
    B4U.G4BW  WSU.G4BW        .00124   WVP.G4BW           61.
    B4U.G4BW  W58.G4BW         127.8
    B5A.G4BW  B5A...BW           -1.   BG4.VOBW            1.
    B5A.G4BW  BG4.WTBW          .684   NGS.G4BW          .684
    B5A.G4BW  WRO.G4BW      92.06722   WMO.G4BW      82.21907
    B5A.G4BW  W12.G4BW          93.8   W02.G4BW          100.
    B5A.G4BW  WVP.G4BW          3.47   W58.G4BW           1.2
    B5B.G4BW  B5B...BW           -1.   BG4.VOBW            1.
    B5B.G4BW  BG4.WTBW          .756   NGS.G4BW          .756
    B5B.G4BW  WRO.G4BW      73.17229   WMO.G4BW      71.98974
    B5B.G4BW  W12.G4BW          92.1   W02.G4BW          100.
    B5B.G4BW  WVP.G4BW          2.36   W58.G4BW         -2.33
    B5C.G4BW  B5C...BW           -1.   BG4.VOBW            1.
    B5C.G4BW  BG4.WTBW          .752   NGS.G4BW          .752
    B5C.G4BW  WRO.G4BW      71.31302   WMO.G4BW      71.11826
    B5C.G4BW  W12.G4BW          99.2   W02.G4BW          100.
    B5C.G4BW  WVP.G4BW          2.71   W58.G4BW           .82
    B5D.G4BW  B5D...BW           -1.   BG4.VOBW            1.
    B5D.G4BW  BG4.WTBW           .75   NGS.G4BW           .75
    B5D.G4BW  WRO.G4BW      76.53985   WMO.G4BW      71.71782
    B5D.G4BW  W12.G4BW          82.2   W02.G4BW          100.
    B5D.G4BW  WVP.G4BW          2.28   W58.G4BW          -2.3
    B5G.G4BW  B5G...BW           -1.   BG4.VOBW            1.
    B5G.G4BW  BG4.WTBW          .666   NGS.G4BW          .666
    B5G.G4BW  WRO.G4BW     102.10471   WMO.G4BW         87.97
    B5G.G4BW  W12.G4BW           91.   W02.G4BW          100.
    B5G.G4BW  WVP.G4BW          3.65   W58.G4BW            .9
    B5H.G4BW  B5H...BW           -1.   BG4.VOBW            1.
    B5H.G4BW  BG4.WTBW          .735   NGS.G4BW          .735
    B5H.G4BW  WRO.G4BW      80.53522   WMO.G4BW      75.95929
    B5H.G4BW  W12.G4BW          85.7   W02.G4BW          100.
    B5H.G4BW  WVP.G4BW          2.55   W58.G4BW          -1.4
    B5I.G4BW  B5I...BW           -1.   BG4.VOBW            1.
    B5I.G4BW  BG4.WTBW          .743   NGS.G4BW          .743
    B5I.G4BW  WRO.G4BW      73.10682   WMO.G4BW      71.95043
    B5I.G4BW  W12.G4BW           98.   W02.G4BW          100.
    B5I.G4BW  WVP.G4BW          2.83   W58.G4BW           1.5
    B5J.G4BW  B5J...BW           -1.   BG4.VOBW            1.
    B5J.G4BW  BG4.WTBW          .715   NGS.G4BW          .715
    B5J.G4BW  WRO.G4BW      92.52513   WMO.G4BW      82.83746
    B5J.G4BW  W12.G4BW          70.2   W02.G4BW          100.
    B5J.G4BW  WVP.G4BW          2.58   W58.G4BW          -1.6
    B5K.G4BW  B5K...BW           -1.   BG4.VOBW            1.
    B5K.G4BW  BG4.WTBW          .684   NGS.G4BW          .684
    B5K.G4BW  WRO.G4BW      92.06722   WMO.G4BW      82.21907
    B5K.G4BW  W12.G4BW          93.8   W02.G4BW          100.
    B5K.G4BW  WVP.G4BW          3.47   W58.G4BW           1.2
    B5L.G4BW  B5L...BW           -1.   BG4.VOBW            1.
    B5L.G4BW  BG4.WTBW          .756   NGS.G4BW          .756
    B5L.G4BW  WRO.G4BW      73.17229   WMO.G4BW      71.98974
    B5L.G4BW  W12.G4BW          92.1   W02.G4BW          100.
    B5L.G4BW  WVP.G4BW          2.36   W58.G4BW         -2.33
    B5M.G4BW  B5M...BW           -1.   BG4.VOBW            1.
    B5M.G4BW  BG4.WTBW          .752   NGS.G4BW          .752
    B5M.G4BW  WRO.G4BW      71.31302   WMO.G4BW      71.11826
    B5M.G4BW  W12.G4BW          99.2   W02.G4BW          100.
    B5M.G4BW  WVP.G4BW          2.71   W58.G4BW           .82
    B5N.G4BW  B5N...BW           -1.   BG4.VOBW            1.
    B5N.G4BW  BG4.WTBW           .75   NGS.G4BW           .75
    B5N.G4BW  WRO.G4BW      76.53985   WMO.G4BW      71.71782
    B5N.G4BW  W12.G4BW          82.2   W02.G4BW          100.
    B5N.G4BW  WVP.G4BW          2.28   W58.G4BW          -2.3
    B5O.G4BW  B5O...BW           -1.   BG4.VOBW            1.
    B5O.G4BW  BG4.WTBW          .753   NGS.G4BW          .753
    B5O.G4BW  WRO.G4BW      74.14662   WMO.G4BW      71.09379
    B5O.G4BW  W12.G4BW          86.2   W02.G4BW          100.
    B5O.G4BW  WVP.G4BW          2.35   W58.G4BW          -1.9
    B5S.G4BW  B5S...BW           -1.   BG4.VOBW            1.
    B5S.G4BW  BG4.WTBW          .625   NGS.G4BW          .625
    B5S.G4BW  WRO.G4BW      93.25986   WMO.G4BW      88.88656
    B5S.G4BW  W12.G4BW          122.   W02.G4BW          100.
    B5S.G4BW  WSU.G4BW        .00156   WVP.G4BW           21.
    B5S.G4BW  W58.G4BW           96.
    BA3.G8BW  BA3...BW           -1.   BG8.VOBW            1.
    BA3.G8BW  BG8.WTBW          .685   NGS.G8BW          .685
    BA3.G8BW  WRO.G8BW      91.85397   WMO.G8BW      87.83771
    BA3.G8BW  W12.G8BW          35.2   W02.G8BW          94.4
    BA3.G8BW  WVP.G8BW            6.   W58.G8BW           6.2
    BA4.G8BW  BA4...BW           -1.   BG8.VOBW            1.
    BA4.G8BW  BG8.WTBW          .703   NGS.G8BW          .703
    BA4.G8BW  WRO.G8BW       95.0259   WMO.G8BW       90.4183
    BA4.G8BW  W12.G8BW          30.2   W02.G8BW           94.
    BA4.G8BW  WVP.G8BW            6.   W58.G8BW           -2.
    BA5.G8BW  BA5...BW           -1.   BG8.VOBW            1.
    BA5.G8BW  BG8.WTBW           .71   NGS.G8BW           .71
    BA5.G8BW  WRO.G8BW      93.37346   WMO.G8BW      89.22463
    BA5.G8BW  W12.G8BW           45.   W02.G8BW           96.
    BA5.G8BW  WVP.G8BW            6.   W58.G8BW          21.1
    BFD.G8BW  BFD...BW           -1.   BG8.VOBW            1.
    BFD.G8BW  BG8.WTBW          .786   NGS.G8BW          .786
    BFD.G8BW  WRO.G8BW      87.33868   WMO.G8BW      78.67732
    BFD.G8BW  W12.G8BW            2.   W02.G8BW           50.
    BFD.G8BW  WSU.G8BW        .09511   WVP.G8BW           2.1
    BFD.G8BW  W58.G8BW           -8.
    BFE.G8BW  BFE...BW           -1.   BG8.VOBW            1.
    BFE.G8BW  BG8.WTBW          .692   NGS.G8BW          .692
    BFE.G8BW  WRO.G8BW      92.03197   WMO.G8BW      81.05526
    BFE.G8BW  W12.G8BW           84.   W02.G8BW          100.
    BFE.G8BW  WSU.G8BW        .03204   WVP.G8BW           10.
    BFE.G8BW  W58.G8BW           47.
    BFF.G8BW  BFF...BW           -1.   BG8.VOBW            1.
    BFF.G8BW  BG8.WTBW           .78   NGS.G8BW           .78
    BFF.G8BW  WRO.G8BW      77.31816   WMO.G8BW      73.78616
    BFF.G8BW  W12.G8BW           10.   W02.G8BW           79.
    BFF.G8BW  WVP.G8BW            4.   W58.G8BW           -2.
    BFH.G8BW  BFH...BW           -1.   BG8.VOBW            1.
    BFH.G8BW  BG8.WTBW          .795   NGS.G8BW          .795
    BFH.G8BW  WRO.G8BW      87.42369   WMO.G8BW      78.28059
    BFH.G8BW  W12.G8BW            1.   W02.G8BW           44.
    BFH.G8BW  WSU.G8BW        .10891   WVP.G8BW            2.
    BFH.G8BW  W58.G8BW           -9.
    BFI.G8BW  BFI...BW           -1.   BG8.VOBW            1.
    BFI.G8BW  BG8.WTBW          .786   NGS.G8BW          .786
    BFI.G8BW  WRO.G8BW      81.33055   WMO.G8BW       76.0993
    BFI.G8BW  W12.G8BW            1.   W02.G8BW           44.
    BFI.G8BW  WSU.G8BW        .00534   WVP.G8BW           2.1
    BFI.G8BW  W58.G8BW           -9.
    BFJ.G8BW  BFJ...BW           -1.   BG8.VOBW            1.
    BFJ.G8BW  BG8.WTBW          .786   NGS.G8BW          .786
    BFJ.G8BW  WRO.G8BW      81.04273   WMO.G8BW      76.21092
    BFJ.G8BW  W12.G8BW            2.   W02.G8BW           50.
    BFJ.G8BW  WSU.G8BW        .00479   WVP.G8BW           2.1
    BFJ.G8BW  W58.G8BW           -8.
    BFK.G8BW  BFK...BW           -1.   BG8.VOBW            1.
    BFK.G8BW  BG8.WTBW          .805   NGS.G8BW          .805
    BFK.G8BW  WRO.G8BW      83.40638   WMO.G8BW      76.79616
    BFK.G8BW  W12.G8BW          -30.   W02.G8BW           20.
    BFK.G8BW  WSU.G8BW        .00064   WVP.G8BW            .5
    BFK.G8BW  W58.G8BW          -21.
    BFL.G8BW  BFL...BW           -1.   BG8.VOBW            1.
    BFL.G8BW  BG8.WTBW          .692   NGS.G8BW          .692
    BFL.G8BW  WRO.G8BW      92.39784   WMO.G8BW       81.5719
    BFL.G8BW  W12.G8BW           84.   W02.G8BW          100.
    BFL.G8BW  WSU.G8BW        .03204   WVP.G8BW           10.
    BFL.G8BW  W58.G8BW           47.
    BFM.G8BW  BFM...BW           -1.   BG8.VOBW            1.
    BFM.G8BW  BG8.WTBW          .692   NGS.G8BW          .692
    BFM.G8BW  WRO.G8BW      92.39784   WMO.G8BW       81.5719
    BFM.G8BW  W12.G8BW           84.   W02.G8BW          100.
    BFM.G8BW  WSU.G8BW        .01605   WVP.G8BW           10.
    BFM.G8BW  W58.G8BW           47.
    BFN.G8BW  BFN...BW           -1.   BG8.VOBW            1.
    BFN.G8BW  BG8.WTBW          .687   NGS.G8BW          .687
    BFN.G8BW  WRO.G8BW      92.70599   WMO.G8BW      81.63251
    BFN.G8BW  W12.G8BW           88.   W02.G8BW          100.
    BFN.G8BW  WSU.G8BW        .00756   WVP.G8BW          12.5
    BFN.G8BW  W58.G8BW           50.
    BFO.G8BW  BFO...BW           -1.   BG8.VOBW            1.
    BFO.G8BW  BG8.WTBW          .795   NGS.G8BW          .795
    BFO.G8BW  WRO.G8BW      85.02961   WMO.G8BW       75.6974
    BFO.G8BW  W12.G8BW            1.   W02.G8BW           44.
    BFO.G8BW  WSU.G8BW        .10891   WVP.G8BW            2.
    BFO.G8BW  W58.G8BW           -9.
    BFP.G8BW  BFP...BW           -1.   BG8.VOBW            1.
    BFP.G8BW  BG8.WTBW          .811   NGS.G8BW          .811
    BFP.G8BW  WRO.G8BW     105.15159   WMO.G8BW      90.22689
    BFP.G8BW  W12.G8BW           2.8   W02.G8BW          79.5
    BFP.G8BW  WVP.G8BW          2.43   W58.G8BW          -3.9
    BFQ.G8BW  BFQ...BW           -1.   BG8.VOBW            1.
    BFQ.G8BW  BG8.WTBW          .811   NGS.G8BW          .811
    BFQ.G8BW  WRO.G8BW     105.15159   WMO.G8BW      90.23941
    BFQ.G8BW  W12.G8BW           2.9   W02.G8BW           80.
    BFQ.G8BW  WVP.G8BW          2.43   W58.G8BW          -3.8
    BFS.G8BW  BFS...BW           -1.   BG8.VOBW            1.
    BFS.G8BW  BG8.WTBW          .687   NGS.G8BW          .687
    BFS.G8BW  WRO.G8BW      92.70599   WMO.G8BW      81.63251
    BFS.G8BW  W12.G8BW           88.   W02.G8BW          100.
    BFS.G8BW  WSU.G8BW        .01889   WVP.G8BW          12.5
    BFS.G8BW  W58.G8BW           50.
    BFT.G8BW  BFT...BW           -1.   BG8.VOBW            1.
    BFT.G8BW  BG8.WTBW           .78   NGS.G8BW           .78
    BFT.G8BW  WRO.G8BW      86.45545   WMO.G8BW      77.40834
    BFT.G8BW  W12.G8BW           10.   W02.G8BW           79.
    BFT.G8BW  WSU.G8BW         .0897   WVP.G8BW            4.
    BFT.G8BW  W58.G8BW           -2.
    BFU.G8BW  BFU...BW           -1.   BG8.VOBW            1.
    BFU.G8BW  BG8.WTBW          .805   NGS.G8BW          .805
    BFU.G8BW  WRO.G8BW      86.74963   WMO.G8BW      77.20759
    BFU.G8BW  W12.G8BW          -30.   W02.G8BW           20.
    BFU.G8BW  WSU.G8BW        .12477   WVP.G8BW            .5
    BFU.G8BW  W58.G8BW          -21.
    BFX.G8BW  BFX...BW           -1.   BG8.VOBW            1.
    BFX.G8BW  BG8.WTBW          .874   NGS.G8BW          .874
    BFX.G8BW  WRO.G8BW     102.56785   WMO.G8BW      87.99456
    BFX.G8BW  W12.G8BW          -31.   W02.G8BW           17.
    BFX.G8BW  WSU.G8BW         .0007   WVP.G8BW            .2
    BFX.G8BW  W58.G8BW          -21.
    BF1.G8BW  BF1...BW           -1.   BG8.VOBW            1.
    BF1.G8BW  BG8.WTBW          .835   NGS.G8BW          .835
    BF1.G8BW  WRO.G8BW     109.60858   WMO.G8BW       92.2351
    BF1.G8BW  W12.G8BW         -14.3   W02.G8BW          74.6
    BF1.G8BW  WVP.G8BW          1.02   W58.G8BW         -13.9
    BF2.G8BW  BF2...BW           -1.   BG8.VOBW            1.
    BF2.G8BW  BG8.WTBW          .736   NGS.G8BW          .736
    BF2.G8BW  WRO.G8BW      82.92617   WMO.G8BW      77.34037
    BF2.G8BW  W12.G8BW          79.6   W02.G8BW          100.
    BF2.G8BW  WVP.G8BW          2.43   W58.G8BW          -1.9
    BF3.G8BW  BF3...BW           -1.   BG8.VOBW            1.
    BF3.G8BW  BG8.WTBW          .835   NGS.G8BW          .835
    BF3.G8BW  WRO.G8BW     109.60858   WMO.G8BW      92.24962
    BF3.G8BW  W12.G8BW         -14.3   W02.G8BW          75.1
    BF3.G8BW  WVP.G8BW          1.02   W58.G8BW         -13.9
    BHC.G8BW  BHC...BW           -1.   BG8.VOBW            1.
    BHC.G8BW  BG8.WTBW          .772   NGS.G8BW          .772
    BHC.G8BW  WRO.G8BW     101.46649   WMO.G8BW      92.87805
    BHC.G8BW  W12.G8BW           22.   W02.G8BW           94.
    BHC.G8BW  WVP.G8BW           3.1   W58.G8BW            1.
    BIS.G8BW  BIS...BW           -1.   BG8.VOBW            1.
    BIS.G8BW  BG8.WTBW          .655   NGS.G8BW          .655
    BIS.G8BW  WRO.G8BW      90.09694   WMO.G8BW      86.13696
    BIS.G8BW  W12.G8BW          108.   W02.G8BW          100.
    BIS.G8BW  WVP.G8BW           13.   W58.G8BW           80.
    BIT.G8BW  BIT...BW           -1.   BG8.VOBW            1.
    BIT.G8BW  BG8.WTBW          .655   NGS.G8BW          .655
    BIT.G8BW  WRO.G8BW      88.88307   WMO.G8BW      84.80651
    BIT.G8BW  W12.G8BW          108.   W02.G8BW          100.
    BIT.G8BW  WVP.G8BW          13.2   W58.G8BW           80.
    BI4.G8BW  BI4...BW           -1.   BG8.VOBW            1.
    BI4.G8BW  BG8.WTBW          .563   NGS.G8BW          .563
    BI4.G8BW  WRO.G8BW     100.92828   WMO.G8BW      94.85789
    BI4.G8BW  W12.G8BW          120.   W02.G8BW          100.
    BI4.G8BW  WSU.G8BW        .00113   WVP.G8BW           78.
    BI4.G8BW  W58.G8BW          134.
    BI5.G8BW  BI5...BW           -1.   BG8.VOBW            1.
    BI5.G8BW  BG8.WTBW          .625   NGS.G8BW          .625
    BI5.G8BW  WRO.G8BW      92.91707   WMO.G8BW      87.95423
    BI5.G8BW  W12.G8BW          115.   W02.G8BW          100.
    BI5.G8BW  WSU.G8BW        .00156   WVP.G8BW           21.
    BI5.G8BW  W58.G8BW           95.
    BLF.G8BW  BLF...BW           -1.   BG8.VOBW            1.
    BLF.G8BW  BG8.WTBW          .658   NGS.G8BW          .658
    BLF.G8BW  WRO.G8BW      75.20602   WMO.G8BW      73.72027
    BLF.G8BW  W12.G8BW          110.   W02.G8BW          101.
    BLF.G8BW  WVP.G8BW          14.8   W58.G8BW           73.
    BLH.G8BW  BLH...BW           -1.   BG8.VOBW            1.
    BLH.G8BW  BG8.WTBW           .71   NGS.G8BW           .71
    BLH.G8BW  WRO.G8BW      70.06795   WMO.G8BW      69.42845
    BLH.G8BW  W12.G8BW           78.   W02.G8BW          100.
    BLH.G8BW  WSU.G8BW         .0213   WVP.G8BW           4.2
    BLH.G8BW  W58.G8BW           25.
    BLI.G8BW  BLI...BW           -1.   BG8.VOBW            1.
    BLI.G8BW  BG8.WTBW          .658   NGS.G8BW          .658
    BLI.G8BW  WRO.G8BW      79.88909   WMO.G8BW      78.12136
    BLI.G8BW  W12.G8BW          110.   W02.G8BW          101.
    BLI.G8BW  WSU.G8BW        .00132   WVP.G8BW          14.5
    BLI.G8BW  W58.G8BW           73.
    BLL.G8BW  BLL...BW           -1.   BG8.VOBW            1.
    BLL.G8BW  BG8.WTBW          .658   NGS.G8BW          .658
    BLL.G8BW  WRO.G8BW      75.20602   WMO.G8BW      73.72027
    BLL.G8BW  W12.G8BW          103.   W02.G8BW          101.
    BLL.G8BW  WSU.G8BW        .01448   WVP.G8BW          14.8
    BLL.G8BW  W58.G8BW           73.
    BLV.G8BW  BLV...BW           -1.   BG8.VOBW            1.
    BLV.G8BW  BG8.WTBW           .67   NGS.G8BW           .67
    BLV.G8BW  WRO.G8BW      74.42509   WMO.G8BW      73.02371
    BLV.G8BW  W12.G8BW           99.   W02.G8BW          103.
    BLV.G8BW  WSU.G8BW        .01675   WVP.G8BW          12.2
    BLV.G8BW  W58.G8BW           59.
    BN4.G8BW  BN4...BW           -1.   BG8.VOBW            1.
    BN4.G8BW  BG8.WTBW          .584   NGS.G8BW          .584
    BN4.G8BW  WRO.G8BW      94.11062   WMO.G8BW      88.35436
    BN4.G8BW  W12.G8BW          119.   W02.G8BW          100.
    BN4.G8BW  WSU.G8BW        .00117   WVP.G8BW           57.
    BN4.G8BW  W58.G8BW          122.
    BN5.G8BW  BN5...BW           -1.   BG8.VOBW            1.
    BN5.G8BW  BG8.WTBW          .631   NGS.G8BW          .631
    BN5.G8BW  WRO.G8BW      71.85843   WMO.G8BW      71.60925
    BN5.G8BW  W12.G8BW          121.   W02.G8BW           99.
    BN5.G8BW  WVP.G8BW           15.   W58.G8BW           86.
    BPE.G8BW  BPE...BW           -1.   BG8.VOBW            1.
    BPE.G8BW  BG8.WTBW          .669   NGS.G8BW          .669
    BPE.G8BW  WRO.G8BW      92.77312   WMO.G8BW      82.21873
    BPE.G8BW  W12.G8BW           98.   W02.G8BW          100.
    BPE.G8BW  WSU.G8BW        .01485   WVP.G8BW          13.5
    BPE.G8BW  W58.G8BW           58.
    BPQ.G8BW  BPQ...BW           -1.   BG8.VOBW            1.
    BPQ.G8BW  BG8.WTBW           .66   NGS.G8BW           .66
    BPQ.G8BW  WRO.G8BW      79.45782   WMO.G8BW      74.74378
    BPQ.G8BW  W12.G8BW          109.   W02.G8BW          100.
    BPQ.G8BW  WVP.G8BW          14.1   W58.G8BW          81.9
    BPT.G8BW  BPT...BW           -1.   BG8.VOBW            1.
    BPT.G8BW  BG8.WTBW          .669   NGS.G8BW          .669
    BPT.G8BW  WRO.G8BW      93.00495   WMO.G8BW      82.29867
    BPT.G8BW  W12.G8BW           98.   W02.G8BW          100.
    BPT.G8BW  WSU.G8BW        .01485   WVP.G8BW          13.5
    BPT.G8BW  W58.G8BW           58.
    BPU.G8BW  BPU...BW           -1.   BG8.VOBW            1.
    BPU.G8BW  BG8.WTBW          .693   NGS.G8BW          .693
    BPU.G8BW  WRO.G8BW      94.18285   WMO.G8BW      79.80062
    BPU.G8BW  W12.G8BW           93.   W02.G8BW          100.
    BPU.G8BW  WSU.G8BW         .0201   WVP.G8BW           13.
    BPU.G8BW  W58.G8BW           50.
    BPV.G8BW  BPV...BW           -1.   BG8.VOBW            1.
    BPV.G8BW  BG8.WTBW          .631   NGS.G8BW          .631
    BPV.G8BW  WRO.G8BW      76.03212   WMO.G8BW      73.15497
    BPV.G8BW  W12.G8BW           93.   W02.G8BW          100.
    BPV.G8BW  WSU.G8BW        .01956   WVP.G8BW           12.
    BPV.G8BW  W58.G8BW           48.
    BTO.G8BW  BTO...BW           -1.   BG8.VOBW            1.
    BTO.G8BW  BG8.WTBW          .872   NGS.G8BW          .872
    BTO.G8BW  WRO.G8BW       104.315   WMO.G8BW      91.00573
    BTO.G8BW  W12.G8BW           18.   W02.G8BW          100.
    BTO.G8BW  WSU.G8BW        .00044   WVP.G8BW           1.6
    BTO.G8BW  W58.G8BW          -20.
    BU4.G8BW  BU4...BW           -1.   BG8.VOBW            1.
    BU4.G8BW  BG8.WTBW           .62   NGS.G8BW           .62
    BU4.G8BW  WRO.G8BW     103.85185   WMO.G8BW      88.31541
    BU4.G8BW  W12.G8BW          113.   W02.G8BW          100.
    BU4.G8BW  WSU.G8BW        .00124   WVP.G8BW           61.
    BU4.G8BW  W58.G8BW         127.8
    BU5.G8BW  BU5...BW           -1.   BG8.VOBW            1.
    BU5.G8BW  BG8.WTBW          .659   NGS.G8BW          .659
    BU5.G8BW  WRO.G8BW     104.38066   WMO.G8BW       84.9222
    BU5.G8BW  W12.G8BW          115.   W02.G8BW          100.
    BU5.G8BW  WVP.G8BW          19.9   W58.G8BW          86.9
    BXA.G8BW  BXA...BW           -1.   BG8.VOBW            1.
    BXA.G8BW  BG8.WTBW          .796   NGS.G8BW          .796
    BXA.G8BW  WRO.G8BW      86.52499   WMO.G8BW      77.67402
    BXA.G8BW  W12.G8BW          -27.   W02.G8BW           28.
    BXA.G8BW  WSU.G8BW        .11303   WVP.G8BW            .7
    BXA.G8BW  W58.G8BW          -19.
    BXB.G8BW  BXB...BW           -1.   BG8.VOBW            1.
    BXB.G8BW  BG8.WTBW          .796   NGS.G8BW          .796
    BXB.G8BW  WRO.G8BW      83.15426   WMO.G8BW      77.10346
    BXB.G8BW  W12.G8BW          -27.   W02.G8BW           28.
    BXB.G8BW  WSU.G8BW        .00056   WVP.G8BW            .7
    BXB.G8BW  W58.G8BW          -19.
    BXC.G8BW  BXC...BW           -1.   BG8.VOBW            1.
    BXC.G8BW  BG8.WTBW          .839   NGS.G8BW          .839
    BXC.G8BW  WRO.G8BW     109.08673   WMO.G8BW      91.58015
    BXC.G8BW  W12.G8BW         -13.2   W02.G8BW          68.1
    BXC.G8BW  WVP.G8BW          1.42   W58.G8BW          -14.
    BXD.G8BW  BXD...BW           -1.   BG8.VOBW            1.
    BXD.G8BW  BG8.WTBW          .786   NGS.G8BW          .786
    BXD.G8BW  WRO.G8BW      85.53494   WMO.G8BW      76.62526
    BXD.G8BW  W12.G8BW            2.   W02.G8BW           50.
    BXD.G8BW  WSU.G8BW        .09511   WVP.G8BW           2.1
    BXD.G8BW  W58.G8BW           -8.
    BXE.G8BW  BXE...BW           -1.   BG8.VOBW            1.
    BXE.G8BW  BG8.WTBW          .723   NGS.G8BW          .723
    BXE.G8BW  WRO.G8BW      86.32131   WMO.G8BW      79.26925
    BXE.G8BW  W12.G8BW           79.   W02.G8BW          100.
    BXE.G8BW  WVP.G8BW          2.63   W58.G8BW           -.9
    BXF.G8BW  BXF...BW           -1.   BG8.VOBW            1.
    BXF.G8BW  BG8.WTBW           .78   NGS.G8BW           .78
    BXF.G8BW  WRO.G8BW      75.96553   WMO.G8BW      72.70699
    BXF.G8BW  W12.G8BW           10.   W02.G8BW           79.
    BXF.G8BW  WVP.G8BW            4.   W58.G8BW           -2.
    BXG.G8BW  BXG...BW           -1.   BG8.VOBW            1.
    BXG.G8BW  BG8.WTBW          .839   NGS.G8BW          .839
    BXG.G8BW  WRO.G8BW     109.08673   WMO.G8BW        91.595
    BXG.G8BW  W12.G8BW         -13.1   W02.G8BW          68.1
    BXG.G8BW  WVP.G8BW          1.42   W58.G8BW          -14.
    BXH.G8BW  BXH...BW           -1.   BG8.VOBW            1.
    BXH.G8BW  BG8.WTBW          .796   NGS.G8BW          .796
    BXH.G8BW  WRO.G8BW      83.73597   WMO.G8BW      74.14184
    BXH.G8BW  W12.G8BW          -27.   W02.G8BW           28.
    BXH.G8BW  WSU.G8BW        .11303   WVP.G8BW            .7
    BXH.G8BW  W58.G8BW          -19.
    BXI.G8BW  BXI...BW           -1.   BG8.VOBW            1.
    BXI.G8BW  BG8.WTBW          .795   NGS.G8BW          .795
    BXI.G8BW  WRO.G8BW      79.08957   WMO.G8BW      73.85299
    BXI.G8BW  W12.G8BW            1.   W02.G8BW           44.
    BXI.G8BW  WSU.G8BW        .00541   WVP.G8BW            2.
    BXI.G8BW  W58.G8BW           -9.
    BXJ.G8BW  BXJ...BW           -1.   BG8.VOBW            1.
    BXJ.G8BW  BG8.WTBW          .786   NGS.G8BW          .786
    BXJ.G8BW  WRO.G8BW      79.34372   WMO.G8BW      74.44831
    BXJ.G8BW  W12.G8BW            2.   W02.G8BW           50.
    BXJ.G8BW  WSU.G8BW        .00479   WVP.G8BW           2.1
    BXJ.G8BW  W58.G8BW           -8.
    BXK.G8BW  BXK...BW           -1.   BG8.VOBW            1.
    BXK.G8BW  BG8.WTBW          .805   NGS.G8BW          .805
    BXK.G8BW  WRO.G8BW      80.33046   WMO.G8BW      73.22415
    BXK.G8BW  W12.G8BW          -30.   W02.G8BW           20.
    BXK.G8BW  WSU.G8BW        .00064   WVP.G8BW            .5
    BXK.G8BW  W58.G8BW          -21.
    BXL.G8BW  BXL...BW           -1.   BG8.VOBW            1.
    BXL.G8BW  BG8.WTBW          .796   NGS.G8BW          .796
    BXL.G8BW  WRO.G8BW      80.50406   WMO.G8BW      73.95196
    BXL.G8BW  W12.G8BW          -27.   W02.G8BW           28.
    BXL.G8BW  WSU.G8BW        .00056   WVP.G8BW            .7
    BXL.G8BW  W58.G8BW          -19.
    BXM.G8BW  BXM...BW           -1.   BG8.VOBW            1.
    BXM.G8BW  BG8.WTBW          .692   NGS.G8BW          .692
    BXM.G8BW  WRO.G8BW      92.03197   WMO.G8BW      81.05526
    BXM.G8BW  W12.G8BW           84.   W02.G8BW          100.
    BXM.G8BW  WSU.G8BW        .01605   WVP.G8BW           10.
    BXM.G8BW  W58.G8BW           47.
    BXN.G8BW  BXN...BW           -1.   BG8.VOBW            1.
    BXN.G8BW  BG8.WTBW          .687   NGS.G8BW          .687
    BXN.G8BW  WRO.G8BW      92.09434   WMO.G8BW      81.45002
    BXN.G8BW  W12.G8BW           88.   W02.G8BW          100.
    BXN.G8BW  WSU.G8BW        .00756   WVP.G8BW          12.5
    BXN.G8BW  W58.G8BW           50.
    BXO.G8BW  BXO...BW           -1.   BG8.VOBW            1.
    BXO.G8BW  BG8.WTBW          .823   NGS.G8BW          .823
    BXO.G8BW  WRO.G8BW     102.34862   WMO.G8BW      88.75606
    BXO.G8BW  W12.G8BW           5.3   W02.G8BW           79.
    BXO.G8BW  WVP.G8BW          2.61   W58.G8BW          -2.8
    BXP.G8BW  BXP...BW           -1.   BG8.VOBW            1.
    BXP.G8BW  BG8.WTBW           .81   NGS.G8BW           .81
    BXP.G8BW  WRO.G8BW     105.15159   WMO.G8BW      90.21435
    BXP.G8BW  W12.G8BW           3.7   W02.G8BW          72.4
    BXP.G8BW  WVP.G8BW          2.71   W58.G8BW          -2.2
    BXQ.G8BW  BXQ...BW           -1.   BG8.VOBW            1.
    BXQ.G8BW  BG8.WTBW           .81   NGS.G8BW           .81
    BXQ.G8BW  WRO.G8BW     105.48297   WMO.G8BW      90.48006
    BXQ.G8BW  W12.G8BW           3.9   W02.G8BW          77.9
    BXQ.G8BW  WVP.G8BW          2.71   W58.G8BW          -2.1
    BXR.G8BW  BXR...BW           -1.   BG8.VOBW            1.
    BXR.G8BW  BG8.WTBW           .81   NGS.G8BW           .81
    BXR.G8BW  WRO.G8BW     105.81583   WMO.G8BW      90.58361
    BXR.G8BW  W12.G8BW            3.   W02.G8BW          77.2
    BXR.G8BW  WVP.G8BW           2.7   W58.G8BW          -2.3
    BXS.G8BW  BXS...BW           -1.   BG8.VOBW            1.
    BXS.G8BW  BG8.WTBW          .687   NGS.G8BW          .687
    BXS.G8BW  WRO.G8BW      92.09434   WMO.G8BW      81.45002
    BXS.G8BW  W12.G8BW           88.   W02.G8BW          100.
    BXS.G8BW  WSU.G8BW        .01889   WVP.G8BW          12.5
    BXS.G8BW  W58.G8BW           50.
    BXT.G8BW  BXT...BW           -1.   BG8.VOBW            1.
    BXT.G8BW  BG8.WTBW           .78   NGS.G8BW           .78
    BXT.G8BW  WRO.G8BW      84.71223   WMO.G8BW       76.0144
    BXT.G8BW  W12.G8BW           10.   W02.G8BW           79.
    BXT.G8BW  WSU.G8BW         .0897   WVP.G8BW            4.
    BXT.G8BW  W58.G8BW           -2.
    BXU.G8BW  BXU...BW           -1.   BG8.VOBW            1.
    BXU.G8BW  BG8.WTBW          .805   NGS.G8BW          .805
    BXU.G8BW  WRO.G8BW      83.48678   WMO.G8BW      73.24984
    BXU.G8BW  W12.G8BW          -30.   W02.G8BW           20.
    BXU.G8BW  WSU.G8BW        .12477   WVP.G8BW            .5
    BXU.G8BW  W58.G8BW          -21.
    BXX.G8BW  BXX...BW           -1.   BG8.VOBW            1.
    BXX.G8BW  BG8.WTBW          .874   NGS.G8BW          .874
    BXX.G8BW  WRO.G8BW     102.56785   WMO.G8BW      87.99456
    BXX.G8BW  W12.G8BW          -31.   W02.G8BW           17.
    BXX.G8BW  WSU.G8BW         .0007   WVP.G8BW            .2
    BXX.G8BW  W58.G8BW          -21.
    BX1.G8BW  BX1...BW           -1.   BG8.VOBW            1.
    BX1.G8BW  BG8.WTBW          .753   NGS.G8BW          .753
    BX1.G8BW  WRO.G8BW      74.14662   WMO.G8BW      71.09379
    BX1.G8BW  W12.G8BW          86.2   W02.G8BW          100.
    BX1.G8BW  WVP.G8BW          2.35   W58.G8BW          -1.9
    BX2.G8BW  BX2...BW           -1.   BG8.VOBW            1.
    BX2.G8BW  BG8.WTBW          .717   NGS.G8BW          .717
    BX2.G8BW  WRO.G8BW       89.0308   WMO.G8BW      80.78009
    BX2.G8BW  W12.G8BW          76.4   W02.G8BW          100.
    BX2.G8BW  WVP.G8BW          2.68   W58.G8BW           -.8
    BX3.G8BW  BX3...BW           -1.   BG8.VOBW            1.
    BX3.G8BW  BG8.WTBW          .852   NGS.G8BW          .852
    BX3.G8BW  WRO.G8BW     105.65125   WMO.G8BW      90.05458
    BX3.G8BW  W12.G8BW         -12.6   W02.G8BW           73.
    BX3.G8BW  WVP.G8BW          1.37   W58.G8BW          -14.
    BX4.G8BW  BX4...BW           -1.   BG8.VOBW            1.
    BX4.G8BW  BG8.WTBW          .838   NGS.G8BW          .838
    BX4.G8BW  WRO.G8BW     109.64842   WMO.G8BW      91.83832
    BX4.G8BW  W12.G8BW         -13.4   W02.G8BW          71.3
    BX4.G8BW  WVP.G8BW          1.48   W58.G8BW          -14.
    BX6.G8BW  BX6...BW           -1.   BG8.VOBW            1.
    BX6.G8BW  BG8.WTBW          .693   NGS.G8BW          .693
    BX6.G8BW  WRO.G8BW      93.74377   WMO.G8BW      79.60009
    BX6.G8BW  W12.G8BW           93.   W02.G8BW          100.
    BX6.G8BW  WSU.G8BW         .0201   WVP.G8BW           13.
    BX6.G8BW  W58.G8BW           50.
    B1A.G8BW  B1A...BW           -1.   BG8.VOBW            1.
    B1A.G8BW  BG8.WTBW          .693   NGS.G8BW          .693
    B1A.G8BW  WRO.G8BW      94.90093   WMO.G8BW      84.96442
    B1A.G8BW  W12.G8BW           71.   W02.G8BW          100.
    B1A.G8BW  WVP.G8BW          5.56   W58.G8BW          23.5
    B1B.G8BW  B1B...BW           -1.   BG8.VOBW            1.
    B1B.G8BW  BG8.WTBW          .806   NGS.G8BW          .806
    B1B.G8BW  WRO.G8BW      99.30716   WMO.G8BW      87.42257
    B1B.G8BW  W12.G8BW         20.77   W02.G8BW           82.
    B1B.G8BW  WVP.G8BW          3.47   W58.G8BW          3.55
    B1C.G8BW  B1C...BW           -1.   BG8.VOBW            1.
    B1C.G8BW  BG8.WTBW          .845   NGS.G8BW          .845
    B1C.G8BW  WRO.G8BW      98.90249   WMO.G8BW      86.84432
    B1C.G8BW  W12.G8BW          2.67   W02.G8BW          22.1
    B1C.G8BW  WVP.G8BW          2.98   W58.G8BW          -2.4
    B1D.G8BW  B1D...BW           -1.   BG8.VOBW            1.
    B1D.G8BW  BG8.WTBW          .823   NGS.G8BW          .823
    B1D.G8BW  WRO.G8BW     103.69118   WMO.G8BW      89.42014
    B1D.G8BW  W12.G8BW           2.9   W02.G8BW          79.8
    B1D.G8BW  WVP.G8BW          2.36   W58.G8BW          -4.4
    B1G.G8BW  B1G...BW           -1.   BG8.VOBW            1.
    B1G.G8BW  BG8.WTBW          .688   NGS.G8BW          .688
    B1G.G8BW  WRO.G8BW      100.1228   WMO.G8BW      88.21309
    B1G.G8BW  W12.G8BW          64.3   W02.G8BW          100.
    B1G.G8BW  WVP.G8BW          5.52   W58.G8BW          22.2
    B1H.G8BW  B1H...BW           -1.   BG8.VOBW            1.
    B1H.G8BW  BG8.WTBW            .8   NGS.G8BW            .8
    B1H.G8BW  WRO.G8BW     102.96806   WMO.G8BW      89.46263
    B1H.G8BW  W12.G8BW          16.5   W02.G8BW          79.8
    B1H.G8BW  WVP.G8BW          3.61   W58.G8BW           4.4
    B1I.G8BW  B1I...BW           -1.   BG8.VOBW            1.
    B1I.G8BW  BG8.WTBW          .841   NGS.G8BW          .841
    B1I.G8BW  WRO.G8BW     101.95318   WMO.G8BW      88.53996
    B1I.G8BW  W12.G8BW           2.9   W02.G8BW          19.9
    B1I.G8BW  WVP.G8BW          3.09   W58.G8BW          -1.6
    B1J.G8BW  B1J...BW           -1.   BG8.VOBW            1.
    B1J.G8BW  BG8.WTBW           .81   NGS.G8BW           .81
    B1J.G8BW  WRO.G8BW     106.96742   WMO.G8BW      91.25379
    B1J.G8BW  W12.G8BW            .7   W02.G8BW          78.8
    B1J.G8BW  WVP.G8BW          2.42   W58.G8BW          -4.1
    B1K.G8BW  B1K...BW           -1.   BG8.VOBW            1.
    B1K.G8BW  BG8.WTBW          .693   NGS.G8BW          .693
    B1K.G8BW  WRO.G8BW      94.90093   WMO.G8BW      84.96442
    B1K.G8BW  W12.G8BW           71.   W02.G8BW          100.
    B1K.G8BW  WVP.G8BW          5.56   W58.G8BW          23.5
    B1L.G8BW  B1L...BW           -1.   BG8.VOBW            1.
    B1L.G8BW  BG8.WTBW          .806   NGS.G8BW          .806
    B1L.G8BW  WRO.G8BW      99.30716   WMO.G8BW      87.42257
    B1L.G8BW  W12.G8BW         20.77   W02.G8BW           82.
    B1L.G8BW  WVP.G8BW          3.47   W58.G8BW          3.55
    B1M.G8BW  B1M...BW           -1.   BG8.VOBW            1.
    B1M.G8BW  BG8.WTBW          .845   NGS.G8BW          .845
    B1M.G8BW  WRO.G8BW      98.90249   WMO.G8BW      86.84432
    B1M.G8BW  W12.G8BW          2.67   W02.G8BW          22.1
    B1M.G8BW  WVP.G8BW          2.98   W58.G8BW          -2.4
    B1N.G8BW  B1N...BW           -1.   BG8.VOBW            1.
    B1N.G8BW  BG8.WTBW          .823   NGS.G8BW          .823
    B1N.G8BW  WRO.G8BW     103.69118   WMO.G8BW      89.42014
    B1N.G8BW  W12.G8BW           2.9   W02.G8BW          79.8
    B1N.G8BW  WVP.G8BW          2.36   W58.G8BW          -4.4
    B1O.G8BW  B1O...BW           -1.   BG8.VOBW            1.
    B1O.G8BW  BG8.WTBW          .823   NGS.G8BW          .823
    B1O.G8BW  WRO.G8BW     102.34862   WMO.G8BW      88.75606
    B1O.G8BW  W12.G8BW           5.3   W02.G8BW           79.
    B1O.G8BW  WVP.G8BW          2.61   W58.G8BW          -2.8
    B2A.G8BW  B2A...BW           -1.   BG8.VOBW            1.
    B2A.G8BW  BG8.WTBW           .74   NGS.G8BW           .74
    B2A.G8BW  WRO.G8BW     108.32812   WMO.G8BW      92.94314
    B2A.G8BW  W12.G8BW          19.7   W02.G8BW          95.2
    B2A.G8BW  WVP.G8BW          1.18   W58.G8BW          -9.4
    B2B.G8BW  B2B...BW           -1.   BG8.VOBW            1.
    B2B.G8BW  BG8.WTBW          .848   NGS.G8BW          .848
    B2B.G8BW  WRO.G8BW     102.76529   WMO.G8BW      88.35055
    B2B.G8BW  W12.G8BW           -7.   W02.G8BW          73.8
    B2B.G8BW  WVP.G8BW          1.17   W58.G8BW         -13.4
    B2C.G8BW  B2C...BW           -1.   BG8.VOBW            1.
    B2C.G8BW  BG8.WTBW          .885   NGS.G8BW          .885
    B2C.G8BW  WRO.G8BW     100.04377   WMO.G8BW      86.52755
    B2C.G8BW  W12.G8BW         -16.7   W02.G8BW           7.3
    B2C.G8BW  WVP.G8BW          1.38   W58.G8BW         -14.7
    B2D.G8BW  B2D...BW           -1.   BG8.VOBW            1.
    B2D.G8BW  BG8.WTBW          .847   NGS.G8BW          .847
    B2D.G8BW  WRO.G8BW     107.43624   WMO.G8BW       91.2576
    B2D.G8BW  W12.G8BW          -15.   W02.G8BW          75.1
    B2D.G8BW  WVP.G8BW          1.02   W58.G8BW         -13.9
    B2G.G8BW  B2G...BW           -1.   BG8.VOBW            1.
    B2G.G8BW  BG8.WTBW          .738   NGS.G8BW          .738
    B2G.G8BW  WRO.G8BW     114.92604   WMO.G8BW      97.63681
    B2G.G8BW  W12.G8BW          14.1   W02.G8BW          100.
    B2G.G8BW  WVP.G8BW          1.13   W58.G8BW          -9.7
    B2H.G8BW  B2H...BW           -1.   BG8.VOBW            1.
    B2H.G8BW  BG8.WTBW          .844   NGS.G8BW          .844
    B2H.G8BW  WRO.G8BW       108.951   WMO.G8BW       92.2581
    B2H.G8BW  W12.G8BW          -9.2   W02.G8BW          71.1
    B2H.G8BW  WVP.G8BW          1.32   W58.G8BW         -13.5
    B2I.G8BW  B2I...BW           -1.   BG8.VOBW            1.
    B2I.G8BW  BG8.WTBW          .883   NGS.G8BW          .883
    B2I.G8BW  WRO.G8BW     103.65702   WMO.G8BW      88.58483
    B2I.G8BW  W12.G8BW         -16.9   W02.G8BW           7.5
    B2I.G8BW  WVP.G8BW          1.35   W58.G8BW         -14.6
    B2J.G8BW  B2J...BW           -1.   BG8.VOBW            1.
    B2J.G8BW  BG8.WTBW          .834   NGS.G8BW          .834
    B2J.G8BW  WRO.G8BW     110.38305   WMO.G8BW      92.88876
    B2J.G8BW  W12.G8BW         -15.9   W02.G8BW          74.3
    B2J.G8BW  WVP.G8BW          1.01   W58.G8BW         -13.9
    B2K.G8BW  B2K...BW           -1.   BG8.VOBW            1.
    B2K.G8BW  BG8.WTBW           .74   NGS.G8BW           .74
    B2K.G8BW  WRO.G8BW     108.32812   WMO.G8BW      92.94314
    B2K.G8BW  W12.G8BW          19.7   W02.G8BW          95.2
    B2K.G8BW  WVP.G8BW          1.18   W58.G8BW          -9.4
    B2L.G8BW  B2L...BW           -1.   BG8.VOBW            1.
    B2L.G8BW  BG8.WTBW          .848   NGS.G8BW          .848
    B2L.G8BW  WRO.G8BW     102.76529   WMO.G8BW      88.35055
    B2L.G8BW  W12.G8BW           -7.   W02.G8BW          73.8
    B2L.G8BW  WVP.G8BW          1.17   W58.G8BW         -13.4
    B2M.G8BW  B2M...BW           -1.   BG8.VOBW            1.
    B2M.G8BW  BG8.WTBW          .885   NGS.G8BW          .885
    B2M.G8BW  WRO.G8BW     100.04377   WMO.G8BW      86.52755
    B2M.G8BW  W12.G8BW         -16.7   W02.G8BW           7.3
    B2M.G8BW  WVP.G8BW          1.38   W58.G8BW         -14.7
    B2N.G8BW  B2N...BW           -1.   BG8.VOBW            1.
    B2N.G8BW  BG8.WTBW          .847   NGS.G8BW          .847
    B2N.G8BW  WRO.G8BW     107.43624   WMO.G8BW       91.2576
    B2N.G8BW  W12.G8BW          -15.   W02.G8BW          75.1
    B2N.G8BW  WVP.G8BW          1.02   W58.G8BW         -13.9
    B2O.G8BW  B2O...BW           -1.   BG8.VOBW            1.
    B2O.G8BW  BG8.WTBW          .852   NGS.G8BW          .852
    B2O.G8BW  WRO.G8BW     105.65125   WMO.G8BW      90.05458
    B2O.G8BW  W12.G8BW         -12.6   W02.G8BW           73.
    B2O.G8BW  WVP.G8BW          1.37   W58.G8BW          -14.
    B4U.G8BW  B4U...BW           -1.   BG8.VOBW            1.
    B4U.G8BW  BG8.WTBW           .62   NGS.G8BW           .62
    B4U.G8BW  WRO.G8BW     103.85185   WMO.G8BW      88.31541
    B4U.G8BW  W12.G8BW          113.   W02.G8BW          100.
    B4U.G8BW  WSU.G8BW        .00124   WVP.G8BW           61.
    B4U.G8BW  W58.G8BW         127.8
    B5A.G8BW  B5A...BW           -1.   BG8.VOBW            1.
    B5A.G8BW  BG8.WTBW          .684   NGS.G8BW          .684
    B5A.G8BW  WRO.G8BW      92.06722   WMO.G8BW      82.21907
    B5A.G8BW  W12.G8BW          93.8   W02.G8BW          100.
    B5A.G8BW  WVP.G8BW          3.47   W58.G8BW           1.2
    B5B.G8BW  B5B...BW           -1.   BG8.VOBW            1.
    B5B.G8BW  BG8.WTBW          .756   NGS.G8BW          .756
    B5B.G8BW  WRO.G8BW      73.17229   WMO.G8BW      71.98974
    B5B.G8BW  W12.G8BW          92.1   W02.G8BW          100.
    B5B.G8BW  WVP.G8BW          2.36   W58.G8BW         -2.33
    B5C.G8BW  B5C...BW           -1.   BG8.VOBW            1.
    B5C.G8BW  BG8.WTBW          .752   NGS.G8BW          .752
    B5C.G8BW  WRO.G8BW      71.31302   WMO.G8BW      71.11826
    B5C.G8BW  W12.G8BW          99.2   W02.G8BW          100.
    B5C.G8BW  WVP.G8BW          2.71   W58.G8BW           .82
    B5D.G8BW  B5D...BW           -1.   BG8.VOBW            1.
    B5D.G8BW  BG8.WTBW           .75   NGS.G8BW           .75
    B5D.G8BW  WRO.G8BW      76.53985   WMO.G8BW      71.71782
    B5D.G8BW  W12.G8BW          82.2   W02.G8BW          100.
    B5D.G8BW  WVP.G8BW          2.28   W58.G8BW          -2.3
    B5G.G8BW  B5G...BW           -1.   BG8.VOBW            1.
    B5G.G8BW  BG8.WTBW          .666   NGS.G8BW          .666
    B5G.G8BW  WRO.G8BW     102.10471   WMO.G8BW         87.97
    B5G.G8BW  W12.G8BW           91.   W02.G8BW          100.
    B5G.G8BW  WVP.G8BW          3.65   W58.G8BW            .9
    B5H.G8BW  B5H...BW           -1.   BG8.VOBW            1.
    B5H.G8BW  BG8.WTBW          .735   NGS.G8BW          .735
    B5H.G8BW  WRO.G8BW      80.53522   WMO.G8BW      75.95929
    B5H.G8BW  W12.G8BW          85.7   W02.G8BW          100.
    B5H.G8BW  WVP.G8BW          2.55   W58.G8BW          -1.4
    B5I.G8BW  B5I...BW           -1.   BG8.VOBW            1.
    B5I.G8BW  BG8.WTBW          .743   NGS.G8BW          .743
    B5I.G8BW  WRO.G8BW      73.10682   WMO.G8BW      71.95043
    B5I.G8BW  W12.G8BW           98.   W02.G8BW          100.
    B5I.G8BW  WVP.G8BW          2.83   W58.G8BW           1.5
    B5J.G8BW  B5J...BW           -1.   BG8.VOBW            1.
    B5J.G8BW  BG8.WTBW          .715   NGS.G8BW          .715
    B5J.G8BW  WRO.G8BW      92.52513   WMO.G8BW      82.83746
    B5J.G8BW  W12.G8BW          70.2   W02.G8BW          100.
    B5J.G8BW  WVP.G8BW          2.58   W58.G8BW          -1.6
    B5K.G8BW  B5K...BW           -1.   BG8.VOBW            1.
    B5K.G8BW  BG8.WTBW          .684   NGS.G8BW          .684
    B5K.G8BW  WRO.G8BW      92.06722   WMO.G8BW      82.21907
    B5K.G8BW  W12.G8BW          93.8   W02.G8BW          100.
    B5K.G8BW  WVP.G8BW          3.47   W58.G8BW           1.2
    B5L.G8BW  B5L...BW           -1.   BG8.VOBW            1.
    B5L.G8BW  BG8.WTBW          .756   NGS.G8BW          .756
    B5L.G8BW  WRO.G8BW      73.17229   WMO.G8BW      71.98974
    B5L.G8BW  W12.G8BW          92.1   W02.G8BW          100.
    B5L.G8BW  WVP.G8BW          2.36   W58.G8BW         -2.33
    B5M.G8BW  B5M...BW           -1.   BG8.VOBW            1.
    B5M.G8BW  BG8.WTBW          .752   NGS.G8BW          .752
    B5M.G8BW  WRO.G8BW      71.31302   WMO.G8BW      71.11826
    B5M.G8BW  W12.G8BW          99.2   W02.G8BW          100.
    B5M.G8BW  WVP.G8BW          2.71   W58.G8BW           .82
    B5N.G8BW  B5N...BW           -1.   BG8.VOBW            1.
    B5N.G8BW  BG8.WTBW           .75   NGS.G8BW           .75
    B5N.G8BW  WRO.G8BW      76.53985   WMO.G8BW      71.71782
    B5N.G8BW  W12.G8BW          82.2   W02.G8BW          100.
    B5N.G8BW  WVP.G8BW          2.28   W58.G8BW          -2.3
    B5O.G8BW  B5O...BW           -1.   BG8.VOBW            1.
    B5O.G8BW  BG8.WTBW          .753   NGS.G8BW          .753
    B5O.G8BW  WRO.G8BW      74.14662   WMO.G8BW      71.09379
    B5O.G8BW  W12.G8BW          86.2   W02.G8BW          100.
    B5O.G8BW  WVP.G8BW          2.35   W58.G8BW          -1.9
    B5S.G8BW  B5S...BW           -1.   BG8.VOBW            1.
    B5S.G8BW  BG8.WTBW          .625   NGS.G8BW          .625
    B5S.G8BW  WRO.G8BW      93.25986   WMO.G8BW      88.88656
    B5S.G8BW  W12.G8BW          122.   W02.G8BW          100.
    B5S.G8BW  WSU.G8BW        .00156   WVP.G8BW           21.
    B5S.G8BW  W58.G8BW           96.
    COF.A.BW  COF.A.BW            1.
    CTF.A.BW  CTF.A.BW            1.
    CTP.A.BW  CTP.A.BW            1.
    CTF.ASBW  CTF.ASBW            1.
    CTF.C.BW  CTF.C.BW            1.
    CBF.F.BW  CBF.F.BW            1.
    CTF.FTBW  CTF.FTBW            1.
    CTF.F2BW  CTF.F2BW            1.
    CTF.LSBW  CTF.LSBW            1.
    CBF.P1BW  CBF.P1BW            1.
    CTC.P1BW  CTC.P1BW            1.
    CTF.P1BW  CTF.P1BW            1.
    CTV.P1BW  CTV.P1BW            1.
    CBF.P2BW  CBF.P2BW            1.
    CTC.P2BW  CTC.P2BW            1.
    CTF.P2BW  CTF.P2BW            1.
    CTV.P2BW  CTV.P2BW            1.
    CTC.P3BW  CTC.P3BW            1.
    CTV.P3BW  CTV.P3BW            1.
    CBF.2XBW  CBF.2XBW            1.
    CTC.2XBW  CTC.2XBW            1.
    CTV.2XBW  CTV.2XBW            1.
    DC3...BW  BC3...BW           -1.
    DC4...BW  BC4...BW           -1.
    DLP...BW  BLP...BW           -1.
    DR1...BW  BR1...BW           -1.
    DYU...BW  BYU...BW           -1.
    D4....BW  B4....BW           -1.
    D4Z...BW  B4Z...BW           -1.
    D6....BW  B6....BW           -1.
    D8....BW  B8....BW           -1.
    D9....BW  B9....BW           -1.
    EBI...BW  BSI...BW            .3   BSJ...BW            .7
    ECD...BW  BCD...BW            1.
    EHC...BW  BHC...BW            1.
    ELI...BW  BLI...BW            1.
    EPI...BW  BPI...BW            1.
    ETO...BW  BTO...BW            1.
    GBF.P2BW  CBF.P2BW          -2.3   CBF.2XBW           3.3
    HBF.P1BW  CBF.P1BW            1.
    HBF.P2BW  CBF.P2BW            1.   U9R...BW          .033
    ITF.E.BW  CTF.E.BW            1.
    ITF.FHBW  CTF.FHBW            1.
    ITF.FMBW  CTF.FMBW            1.
    ITF.FSBW  CTF.FSBW            1.
    ITP.FTBW  CTP.FTBW            1.
    ITF.H.BW  CTF.H.BW            1.
    ITF.I.BW  CTF.I.BW            1.
    ITF.IHBW  CTF.IHBW            1.
    ITF.ISBW  CTF.ISBW            1.
    ITF.PHBW  CTF.PHBW            1.
    ITF.PSBW  CTF.PSBW            1.
    ITF.P3BW  CTF.P3BW            1.
    MT1.R1BW  BR1...BW            1.   BG2...BW          -.71
    MT1.R1BW  BG4...BW          -.15   BG8...BW          -.14
    PSK.A.BW  UFU...BW         -.135   UNF...BW          -.16
    PSK.A.BW  U.S...BW        -.0231   U9R...BW        -.0011
    PSK.A.BW  CTF.A.BW        -2.196   CTP.A.BW        -1.816
    PSK.A.BW  COF.A.BW           -1.   BSK...BW           -1.
    PSK.A.BW  BA4...BW         1.816   BSI...BW         -1.14
    PSK.A.BW  BSJ...BW         -.056
    P3U.A.BW  UFU...BW         -.128   UNF...BW          -.16
    P3U.A.BW  U.S...BW        -.0315   U9R...BW        -.0011
    P3U.A.BW  CTF.A.BW        -2.198   CTP.A.BW        -1.727
    P3U.A.BW  COF.A.BW           -1.   B3U...BW           -1.
    P3U.A.BW  BA3...BW         1.727   BSI...BW        -1.181
    P3U.A.BW  BSJ...BW         -.017
    P5U.A.BW  UFU...BW         -.148   UNF...BW          -.16
    P5U.A.BW  U.S...BW        -.0231   U9R...BW        -.0011
    P5U.A.BW  CTF.A.BW        -2.295   CTP.A.BW           -2.
    P5U.A.BW  COF.A.BW           -1.   B5U...BW           -1.
    P5U.A.BW  BA5...BW            2.   BSI...BW        -1.273
    P5U.A.BW  BSJ...BW         -.022
    PTS.ASBW  UFU...BW         -.007   CTF.ASBW           -1.
    PTS.ASBW  B5U...BW          .195   B5S...BW            .1
    PTS.ASBW  BPV...BW          .705   BTS...BW           -1.
    PTT.ASBW  UFU...BW         -.007   CTF.ASBW           -1.
    PTT.ASBW  B5U...BW           .18   BTT...BW           -1.
    PTT.ASBW  B5S...BW            .1   BPV...BW           .72
    PTU.ASBW  UFU...BW         -.007   CTF.ASBW           -1.
    PTU.ASBW  B5U...BW           .23   BTU...BW           -1.
    PTU.ASBW  B5S...BW           .05   BPV...BW           .72
    PTV.ASBW  UFU...BW         -.007   CTF.ASBW           -1.
    PTV.ASBW  B5U...BW           .25   B5S...BW           .05
    PTV.ASBW  BPV...BW            .7   BTV...BW           -1.
    PCD.C.BW  UFU...BW        -.0016   UH2...BW       -.11501
    PCD.C.BW  UNF...BW        -.0016   U9R...BW        -.0155
    PCD.C.BW  CTF.C.BW           -1.   BSI...BW          .004
    PCD.C.BW  BSQ...BW         .0727   BSO...BW          .126
    PCD.C.BW  BSP...BW         .0183   BCP...BW          .348
    PCD.C.BW  BCD...BW           -1.   BSJ...BW         .0129
    PCD.C.BW  BH....BW       -.00585
    PSA.E.BW  UFU...BW         -.006   UNF...BW         -.033
    PSA.E.BW  U9R...BW        -.0292   CTF.E.BW           -1.
    PSA.E.BW  BFY...BW           .51   BFX...BW           .49
    PSA.E.BW  BSA...BW           -1.
    PSB.E.BW  UFU...BW         -.006   UNF...BW         -.033
    PSB.E.BW  U9R...BW        -.0292   CTF.E.BW           -1.
    PSB.E.BW  BFY...BW           .42   BFX...BW           .58
    PSB.E.BW  BSB...BW           -1.
    PSC.E.BW  UFU...BW         -.006   UNF...BW         -.033
    PSC.E.BW  U9R...BW        -.0292   CTF.E.BW           -1.
    PSC.E.BW  BSC...BW           -1.   BFY...BW           .45
    PSC.E.BW  BFX...BW           .55
    PSD.E.BW  UFU...BW         -.006   UNF...BW         -.033
    PSD.E.BW  U9R...BW        -.0292   CTF.E.BW           -1.
    PSD.E.BW  BFY...BW           .35   BFX...BW           .65
    PSD.E.BW  BSD...BW           -1.
    PCP.F.BW  URX...BW        -1.366   UH2...BW       -.08513
    PCP.F.BW  UFU...BW          .005   U9R...BW        -.0068
    PCP.F.BW  CTF.F2BW           -1.   CBF.F.BW           -1.
    PCP.F.BW  BTT...BW          .057   BTL...BW          .172
    PCP.F.BW  B8....BW          .238   BYU...BW          .034
    PCP.F.BW  BTH...BW          .344   BU4...BW         .0064
    PCP.F.BW  B5S...BW         .0103   BCP...BW           -1.
    PCP.F.BW  B3U...BW         .0564   BSI...BW         .0408
    PCP.F.BW  BSJ...BW          .012   BSK...BW          .066
    PCP.F.BW  BGS...BW         .0435   BLP...BW         .0234
    PCP.F.BW  BH....BW       -.00433   BU5...BW         .0067
    PCP.F.BW  BN5...BW         .0048
    PPI.F.BW  URX...BW        -1.366   UH2...BW       -.08513
    PPI.F.BW  UFU...BW          .005   U9R...BW        -.0068
    PPI.F.BW  CTF.F2BW           -1.   CBF.F.BW           -1.
    PPI.F.BW  BTT...BW          .057   BTL...BW          .172
    PPI.F.BW  B8....BW          .238   BYU...BW          .034
    PPI.F.BW  BTH...BW          .344   BU4...BW         .0064
    PPI.F.BW  B5S...BW         .0103   BPI...BW           -1.
    PPI.F.BW  B3U...BW         .0564   BSI...BW         .0408
    PPI.F.BW  BSJ...BW          .012   BSK...BW          .066
    PPI.F.BW  BGS...BW         .0435   BLP...BW         .0234
    PPI.F.BW  BH....BW       -.00433   BU5...BW         .0067
    PPI.F.BW  BN5...BW         .0048
    PTH.FDBW  BTJ...BW          .901   B8....BW          .099
    PTH.FDBW  BTH...BW           -1.
    PTI.FDBW  B8....BW          .099   BTI...BW           -1.
    PTI.FDBW  BTK...BW          .901
    PSE.FHBW  UH2...BW       -.21331   UFU...BW         -.002
    PSE.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PSE.FHBW  BFJ...BW            1.   BH....BW       -.01085
    PSE.FHBW  BSE...BW           -1.
    PSF.FHBW  UH2...BW       -.21331   UFU...BW         -.002
    PSF.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PSF.FHBW  BSF...BW           -1.   BXJ...BW            1.
    PSF.FHBW  BH....BW       -.01085
    PSG.FHBW  UH2...BW       -.21508   UFU...BW         -.002
    PSG.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PSG.FHBW  BSG...BW           -1.   BFI...BW            1.
    PSG.FHBW  BH....BW       -.01094
    PSH.FHBW  UH2...BW       -.21154   UFU...BW         -.002
    PSH.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PSH.FHBW  BSH...BW           -1.   BXI...BW            1.
    PSH.FHBW  BH....BW       -.01076
    PTN.FHBW  UH2...BW       -.14273   UFU...BW         -.002
    PTN.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PTN.FHBW  BTN...BW           -1.   BSA...BW            1.
    PTN.FHBW  BH....BW       -.00726
    PTR.FHBW  UH2...BW       -.13919   UFU...BW         -.002
    PTR.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PTR.FHBW  BTR...BW           -1.   BSB...BW            1.
    PTR.FHBW  BH....BW       -.00708
    PTW.FHBW  UH2...BW       -.14293   UFU...BW         -.002
    PTW.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PTW.FHBW  BTW...BW           -1.   BSC...BW            1.
    PTW.FHBW  BH....BW       -.00727
    PTZ.FHBW  UH2...BW       -.13939   UFU...BW         -.002
    PTZ.FHBW  U9R...BW         -.002   CTF.FHBW           -1.
    PTZ.FHBW  BTZ...BW           -1.   BSD...BW            1.
    PTZ.FHBW  BH....BW       -.00709
    PTL.FMBW  UNF...BW         -.006   CTF.FMBW           -1.
    PTL.FMBW  BTL...BW           -1.   BFM...BW            1.
    PTM.FMBW  UNF...BW         -.006   CTF.FMBW           -1.
    PTM.FMBW  BTM...BW           -1.   BXM...BW            1.
    PTX.FMBW  UNF...BW         -.006   CTF.FMBW           -1.
    PTX.FMBW  BTX...BW           -1.   BFN...BW            1.
    PTY.FMBW  UNF...BW         -.006   CTF.FMBW           -1.
    PTY.FMBW  BTY...BW           -1.   BXN...BW            1.
    PSE.FSBW  UFU...BW         -.009   CTF.FSBW           -1.
    PSE.FSBW  BTX...BW           .21   BTF...BW          .474
    PSE.FSBW  BTN...BW          .316   BSE...BW           -1.
    PSF.FSBW  UFU...BW         -.009   CTF.FSBW           -1.
    PSF.FSBW  BTR...BW          .316   BTY...BW           .21
    PSF.FSBW  BSF...BW           -1.   BTG...BW          .474
    PSG.FSBW  UFU...BW         -.009   CTF.FSBW           -1.
    PSG.FSBW  BTX...BW            .2   BSG...BW           -1.
    PSG.FSBW  BTF...BW           .45   BTW...BW           .35
    PSH.FSBW  UFU...BW         -.009   CTF.FSBW           -1.
    PSH.FSBW  BTY...BW            .2   BSH...BW           -1.
    PSH.FSBW  BTG...BW           .45   BTZ...BW           .35
    PTL.FSBW  UFU...BW         -.009   CTF.FSBW           -1.
    PTL.FSBW  BTX...BW            .8   BTL...BW           -1.
    PTL.FSBW  BTF...BW            .2
    PTM.FSBW  UFU...BW         -.009   CTF.FSBW           -1.
    PTM.FSBW  BTY...BW            .8   BTM...BW           -1.
    PTM.FSBW  BTG...BW            .2
    PTS.FTBW  UFU...BW         -.009   CTF.FTBW           -1.
    PTS.FTBW  CTP.FTBW          -.38   BI5...BW           .38
    PTS.FTBW  BTV...BW           .62   BTS...BW           -1.
    PTT.FTBW  UFU...BW         -.009   CTF.FTBW           -1.
    PTT.FTBW  CTP.FTBW           -.4   BTT...BW           -1.
    PTT.FTBW  BTU...BW            .6   BI5...BW            .4
    PCP.F1BW  URX...BW        -1.366   UH2...BW       -.08513
    PCP.F1BW  UFU...BW          .005   U9R...BW         -.007
    PCP.F1BW  CTF.F2BW           -1.   BTT...BW       -.16286
    PCP.F1BW  BTL...BW       -.49143   B8....BW          .404
    PCP.F1BW  BYU...BW         .0745   BTH...BW       -.98286
    PCP.F1BW  BU4...BW         .0009   B5S...BW         .0104
    PCP.F1BW  BCP...BW           -1.   BTS...BW         .2159
    PCP.F1BW  BTM...BW         .6504   BTI...BW        1.2999
    PCP.F1BW  B3U...BW        -.0092   BSI...BW         .0262
    PCP.F1BW  BSJ...BW         .0061   BSK...BW         .0074
    PCP.F1BW  BGS...BW        -.0034   BLP...BW         .0047
    PCP.F1BW  BH....BW       -.00433   BU5...BW         .0035
    PCP.F1BW  BN5...BW         .0034
    PPI.F1BW  URX...BW        -1.366   UH2...BW       -.08513
    PPI.F1BW  UFU...BW          .005   U9R...BW         -.007
    PPI.F1BW  CTF.F2BW           -1.   BTT...BW       -.16286
    PPI.F1BW  BTL...BW       -.49143   B8....BW          .404
    PPI.F1BW  BYU...BW         .0745   BTH...BW       -.98286
    PPI.F1BW  BU4...BW         .0009   B5S...BW         .0104
    PPI.F1BW  BTS...BW         .2159   BTM...BW         .6504
    PPI.F1BW  BTI...BW        1.2999   BPI...BW           -1.
    PPI.F1BW  B3U...BW        -.0092   BSI...BW         .0262
    PPI.F1BW  BSJ...BW         .0061   BSK...BW         .0074
    PPI.F1BW  BGS...BW        -.0034   BLP...BW         .0047
    PPI.F1BW  BH....BW       -.00433   BU5...BW         .0035
    PPI.F1BW  BN5...BW         .0034
    PGS.H.BW  UH2...BW         19.66   UFU...BW         -.005
    PGS.H.BW  UNF...BW         -1.99   U9R...BW         -.035
    PGS.H.BW  CTF.H.BW           -1.   BH....BW            1.
    PGS.H.BW  BGS...BW           -1.
    PST.I.BW  UH2...BW        -.2084   UFU...BW         -.005
    PST.I.BW  UNF...BW          -.04   U9R...BW         -.035
    PST.I.BW  CTF.I.BW           -1.   BST...BW           -1.
    PST.I.BW  BIS...BW         .9845   BH....BW        -.0106
    PST.I.BW  BGS...BW          .026
    PSU.I.BW  UH2...BW       -.04325   UFU...BW         -.005
    PSU.I.BW  UNF...BW          -.04   U9R...BW         -.035
    PSU.I.BW  CTF.I.BW           -1.   BIT...BW          .996
    PSU.I.BW  BSU...BW           -1.   BH....BW        -.0022
    PSU.I.BW  BGS...BW         .0075
    PSR.IHBW  UH2...BW       -.00197   UFU...BW         -.002
    PSR.IHBW  U9R...BW         -.009   CTF.IHBW           -1.
    PSR.IHBW  BSR...BW           -1.   BST...BW            1.
    PSR.IHBW  BH....BW        -.0001
    PLI.ISBW  UFU...BW         -.009   CTF.ISBW           -1.
    PLI.ISBW  BLI...BW           -1.   BSR...BW            1.
    PSQ.ISBW  UFU...BW         -.009   CTF.ISBW           -1.
    PSQ.ISBW  BSR...BW          .835   BSN...BW          .165
    PSQ.ISBW  BSQ...BW           -1.
    PSQ.LSBW  UFU...BW         -.009   CTF.LSBW           -1.
    PSQ.LSBW  BSQ...BW           -1.   BSN...BW          .165
    PSQ.LSBW  BK6...BW          .835
    PFY.M.BW  B8....BW            1.   BFY...BW           -1.
    PK6.M.BW  B4....BW            1.   BK6...BW           -1.
    PPZ.M.BW  B8....BW            1.   BPZ...BW           -1.
    PSN.M.BW  BSN...BW           -1.   B4....BW            1.
    PSP.M.BW  B6....BW            1.   BSP...BW           -1.
    PSQ.M.BW  B4....BW            1.   BSQ...BW           -1.
    PSR.M.BW  BSR...BW           -1.   B4....BW            1.
    PSU.M.BW  BSU...BW           -1.   B4....BW            1.
    PTN.M.BW  BTN...BW           -1.   B8....BW            1.
    PTR.M.BW  BTR...BW           -1.   B8....BW            1.
    PTW.M.BW  BTW...BW           -1.   B8....BW            1.
    PTZ.M.BW  BTZ...BW           -1.   B8....BW            1.
    P3U.M.BW  B3U...BW           -1.   BLP...BW            1.
    PSK.M1BW  BSK...BW           -1.   BC4...BW            1.
    P3U.M1BW  B3U...BW           -1.   BC3...BW            1.
    PSI.M2BW  BSI...BW           -1.   B4Z...BW            1.
    PSJ.M2BW  BSJ...BW           -1.   B4Z...BW            1.
    PTF.PHBW  UFU...BW         -.001   UH2...BW       -.31456
    PTF.PHBW  UNF...BW         -.001   U9R...BW         -.005
    PTF.PHBW  CTF.PHBW           -1.   BTF...BW           -1.
    PTF.PHBW  BTD...BW            1.   BH....BW         -.016
    PTG.PHBW  UFU...BW         -.001   UH2...BW       -.31456
    PTG.PHBW  UNF...BW         -.001   U9R...BW         -.005
    PTG.PHBW  CTF.PHBW           -1.   BTG...BW           -1.
    PTG.PHBW  BTE...BW            1.   BH....BW         -.016
    PTP.PSBW  CTF.PSBW           -1.   BSU...BW        .11453
    PTP.PSBW  BF2...BW        .03174   BTB...BW        .85373
    PTP.PSBW  BTP...BW           -1.
    PTQ.PSBW  CTF.PSBW           -1.   BSU...BW         .0986
    PTQ.PSBW  BXE...BW        .03543   BTC...BW        .86597
    PTQ.PSBW  BTQ...BW           -1.
    PXO.PSBW  CTF.PSBW           -1.   BSU...BW        .09499
    PXO.PSBW  BX1...BW        .03912   BX3...BW        .86589
    PXO.PSBW  BXO...BW           -1.
    PXR.PSBW  CTF.PSBW           -1.   BSU...BW        .09414
    PXR.PSBW  BX2...BW        .02672   BX4...BW        .87914
    PXR.PSBW  BXR...BW           -1.
    P1A.PSBW  CTF.PSBW           -1.   B1A...BW           -1.
    P1A.PSBW  BSU...BW        .41363   B5A...BW        .19751
    P1A.PSBW  B2A...BW        .38886
    P1B.PSBW  CTF.PSBW           -1.   B1B...BW           -1.
    P1B.PSBW  BSU...BW        .17016   B5B...BW        .08113
    P1B.PSBW  B2B...BW        .74871
    P1C.PSBW  CTF.PSBW           -1.   B1C...BW           -1.
    P1C.PSBW  BSU...BW        .12236   B5C...BW        .03406
    P1C.PSBW  B2C...BW        .84358
    P1D.PSBW  CTF.PSBW           -1.   B1D...BW           -1.
    P1D.PSBW  BSU...BW        .11126   B5D...BW        .04345
    P1D.PSBW  B2D...BW        .84529
    P1G.PSBW  CTF.PSBW           -1.   B1G...BW           -1.
    P1G.PSBW  BSU...BW        .38745   B5G...BW        .18143
    P1G.PSBW  B2G...BW        .43112
    P1H.PSBW  CTF.PSBW           -1.   B1H...BW           -1.
    P1H.PSBW  BSU...BW        .17429   B5H...BW        .05413
    P1H.PSBW  B2H...BW        .77158
    P1I.PSBW  CTF.PSBW           -1.   BSU...BW        .12875
    P1I.PSBW  B5I...BW        .03067   B2I...BW        .84058
    P1I.PSBW  B1I...BW           -1.
    P1J.PSBW  CTF.PSBW           -1.   BSU...BW        .11263
    P1J.PSBW  B5J...BW        .02902   B2J...BW        .85835
    P1J.PSBW  B1J...BW           -1.
    P1K.PSBW  CTF.PSBW           -1.   BSU...BW        .41363
    P1K.PSBW  B2K...BW        .38886   B1K...BW           -1.
    P1K.PSBW  B5K...BW        .19751
    P1L.PSBW  CTF.PSBW           -1.   BSU...BW        .17016
    P1L.PSBW  B2L...BW        .74871   B1L...BW           -1.
    P1L.PSBW  B5L...BW        .08113
    P1M.PSBW  CTF.PSBW           -1.   BSU...BW        .12236
    P1M.PSBW  B2M...BW        .84358   B1M...BW           -1.
    P1M.PSBW  B5M...BW        .03406
    P1N.PSBW  CTF.PSBW           -1.   BSU...BW        .11126
    P1N.PSBW  B2N...BW        .84529   B1N...BW           -1.
    P1N.PSBW  B5N...BW        .04345
    P1O.PSBW  CTF.PSBW           -1.   BSU...BW        .09499
    P1O.PSBW  B2O...BW        .86589   B1O...BW           -1.
    P1O.PSBW  B5O...BW        .03912
    PPG.P1BW  UFU...BW         -.002   UH2...BW       1.04473
    PPG.P1BW  UNF...BW         -.003   U9R...BW         -.073
    PPG.P1BW  CTF.P1BW           -1.   CBF.P1BW           -1.
    PPG.P1BW  CTC.P1BW       -.00001   CTV.P1BW           -1.
    PPG.P1BW  BPG...BW           -1.   B1B...BW         .7563
    PPG.P1BW  BGS...BW        .05521   BLP...BW        .06964
    PPG.P1BW  BSI...BW        .03369   BSJ...BW        .04936
    PPG.P1BW  BH....BW        .05314
    PPH.P1BW  UFU...BW         -.002   UH2...BW       1.07383
    PPH.P1BW  UNF...BW         -.003   U9R...BW         -.073
    PPH.P1BW  CTF.P1BW           -1.   CBF.P1BW           -1.
    PPH.P1BW  CTC.P1BW       -.00001   CTV.P1BW           -1.
    PPH.P1BW  BPH...BW           -1.   B1C...BW         .8099
    PPH.P1BW  BGS...BW        .04572   BLP...BW        .05646
    PPH.P1BW  BSI...BW        .02506   BSJ...BW        .03991
    PPH.P1BW  BH....BW        .05462
    PTA.P1BW  UFU...BW         -.002   UH2...BW        .98457
    PTA.P1BW  UNF...BW         -.003   U9R...BW         -.073
    PTA.P1BW  CTF.P1BW           -1.   CBF.P1BW           -1.
    PTA.P1BW  CTC.P1BW       -.00001   CTV.P1BW           -1.
    PTA.P1BW  BTA...BW           -1.   B1A...BW          .674
    PTA.P1BW  BGS...BW        .07564   BLP...BW        .10016
    PTA.P1BW  BSI...BW        .05849   BSJ...BW        .07667
    PTA.P1BW  BH....BW        .05008
    PTD.P1BW  UFU...BW         -.002   UH2...BW        .75848
    PTD.P1BW  UNF...BW         -.003   U9R...BW         -.073
    PTD.P1BW  CTF.P1BW           -1.   CBF.P1BW           -1.
    PTD.P1BW  CTC.P1BW           -1.   CTV.P1BW       -.00001
    PTD.P1BW  BTD...BW           -1.   BGS...BW        .04265
    PTD.P1BW  BLP...BW        .05193   BSI...BW        .02378
    PTD.P1BW  BSJ...BW        .03703   B1D...BW          .825
    PTD.P1BW  BH....BW        .03858
    PTE.P1BW  UFU...BW         -.002   UH2...BW        .66746
    PTE.P1BW  UNF...BW         -.003   U9R...BW         -.073
    PTE.P1BW  CTF.P1BW           -1.   CBF.P1BW           -1.
    PTE.P1BW  CTC.P1BW           -1.   CTV.P1BW       -.00001
    PTE.P1BW  BXO...BW         .8455   BGS...BW        .03823
    PTE.P1BW  BLP...BW        .04606   BSI...BW          .021
    PTE.P1BW  BSJ...BW        .03277   BTE...BW           -1.
    PTE.P1BW  BH....BW        .03395
    PZF.P2BW  UFU...BW         -.011   UH2...BW       1.13989
    PZF.P2BW  UNF...BW         -.007   U9R...BW         -.083
    PZF.P2BW  CTF.P2BW           -1.   CBF.P2BW           -1.
    PZF.P2BW  CTC.P2BW       -.00001   CTV.P2BW           -1.
    PZF.P2BW  BZF...BW           -1.   B1G...BW         .6246
    PZF.P2BW  BGS...BW        .08671   BLP...BW        .11562
    PZF.P2BW  BSI...BW          .059   BSJ...BW        .08763
    PZF.P2BW  BH....BW        .05798
    PZG.P2BW  UFU...BW         -.011   UH2...BW       1.19081
    PZG.P2BW  UNF...BW         -.007   U9R...BW         -.083
    PZG.P2BW  CTF.P2BW           -1.   CBF.P2BW           -1.
    PZG.P2BW  CTC.P2BW       -.00001   CTV.P2BW           -1.
    PZG.P2BW  BZG...BW           -1.   B1H...BW         .7186
    PZG.P2BW  BGS...BW        .06316   BLP...BW        .07957
    PZG.P2BW  BSI...BW         .0369   BSJ...BW        .05632
    PZG.P2BW  BH....BW        .06057
    PZH.P2BW  UFU...BW         -.011   UH2...BW       1.19474
    PZH.P2BW  UNF...BW         -.007   U9R...BW         -.083
    PZH.P2BW  CTF.P2BW           -1.   CBF.P2BW           -1.
    PZH.P2BW  CTC.P2BW       -.00001   CTV.P2BW           -1.
    PZH.P2BW  BZH...BW           -1.   B1I...BW         .7833
    PZH.P2BW  BGS...BW        .05104   BLP...BW        .06291
    PZH.P2BW  BSI...BW        .02769   BSJ...BW        .04405
    PZH.P2BW  BH....BW        .06077
    PZI.P2BW  UFU...BW         -.011   UH2...BW        .87153
    PZI.P2BW  UNF...BW         -.007   U9R...BW         -.083
    PZI.P2BW  CTF.P2BW           -1.   CBF.P2BW           -1.
    PZI.P2BW  CTC.P2BW           -1.   CTV.P2BW       -.00001
    PZI.P2BW  BZI...BW           -1.   BGS...BW        .05424
    PZI.P2BW  BLP...BW        .06316   BSI...BW        .02702
    PZI.P2BW  BSJ...BW        .04395   B1J...BW         .7851
    PZI.P2BW  BH....BW        .04438
    PZJ.P2BW  UFU...BW         -.011   UH2...BW        .75848
    PZJ.P2BW  UNF...BW         -.007   U9R...BW         -.083
    PZJ.P2BW  CTF.P2BW           -1.   CBF.P2BW           -1.
    PZJ.P2BW  CTC.P2BW           -1.   CTV.P2BW       -.00001
    PZJ.P2BW  BGS...BW        .04858   BLP...BW         .0554
    PZJ.P2BW  BSI...BW        .02357   BSJ...BW        .03841
    PZJ.P2BW  BXR...BW         .8117   BZJ...BW           -1.
    PZJ.P2BW  BH....BW        .03858
    PTD.P3BW  UFU...BW         -.008   UH2...BW        .86504
    PTD.P3BW  UNF...BW         -.006   U9R...BW          -.04
    PTD.P3BW  CTF.P3BW           -1.   CTC.P3BW           -1.
    PTD.P3BW  CTV.P3BW       -.00001   BTD...BW           -1.
    PTD.P3BW  BTP...BW         .7921   BGS...BW         .0519
    PTD.P3BW  BLP...BW         .0608   BSI...BW         .0263
    PTD.P3BW  BSJ...BW         .0425   BH....BW          .044
    PTE.P3BW  UFU...BW         -.008   UH2...BW        .72349
    PTE.P3BW  UNF...BW         -.006   U9R...BW          -.04
    PTE.P3BW  CTF.P3BW           -1.   CTC.P3BW           -1.
    PTE.P3BW  CTV.P3BW       -.00001   BTQ...BW         .8292
    PTE.P3BW  BGS...BW         .0434   BLP...BW         .0505
    PTE.P3BW  BSI...BW         .0221   BSJ...BW         .0355
    PTE.P3BW  BTE...BW           -1.   BH....BW         .0368
    PK1.TABW  BK1...BW           -1.   B1K...BW            1.
    PK2.TABW  BK2...BW           -1.   B1L...BW            1.
    PK3.TABW  BK3...BW           -1.   B1M...BW            1.
    PK4.TABW  BK4...BW           -1.   B1N...BW            1.
    PK5.TABW  BK5...BW           -1.   B1O...BW            1.
    PTB.TBBW  BTB...BW           -1.   BF1...BW            1.
    PTC.TBBW  BTC...BW           -1.   BXC...BW            1.
    PSE.TDBW  BFD...BW            1.   BSE...BW           -1.
    PSF.TDBW  BXD...BW            1.   BSF...BW           -1.
    PSG.TDBW  BSG...BW           -1.   BFH...BW            1.
    PSH.TDBW  BSH...BW           -1.   BFO...BW            1.
    PTH.TEBW  BTH...BW           -1.   BSG...BW            1.
    PTI.TEBW  BTI...BW           -1.   BSH...BW            1.
    PTJ.TEBW  BTJ...BW           -1.   BSE...BW            1.
    PTK.TEBW  BTK...BW           -1.   BSF...BW            1.
    PTF.TFBW  BTF...BW           -1.   BFT...BW            1.
    PTG.TFBW  BTG...BW           -1.   BXT...BW            1.
    PTD.TGBW  BTD...BW           -1.   BFF...BW            1.
    PTE.TGBW  BTE...BW           -1.   BXF...BW            1.
    PSI.THBW  BI4...BW            1.   BSI...BW           -1.
    PSJ.THBW  BN4...BW            1.   BSJ...BW           -1.
    PSK.THBW  B4U...BW            1.   BSK...BW           -1.
    PTP.TIBW  BTP...BW           -1.   BFP...BW            1.
    PTQ.TIBW  BTQ...BW           -1.   BXP...BW            1.
    PSN.TJBW  BSN...BW           -1.   BTA...BW            1.
    PSO.TJBW  BSO...BW           -1.   BPG...BW            1.
    PSP.TJBW  BSP...BW           -1.   BPH...BW            1.
    PK6.TKBW  BLL...BW            1.   BK6...BW           -1.
    PSN.TKBW  BSN...BW           -1.   BLH...BW            1.
    PSQ.TKBW  BLV...BW            1.   BSQ...BW           -1.
    PSR.TKBW  BLL...BW            1.   BSR...BW           -1.
    PST.TKBW  BLF...BW            1.   BST...BW           -1.
    PSU.TKBW  BSU...BW           -1.   BPQ...BW            1.
    PTL.TLBW  BTL...BW           -1.   BFL...BW            1.
    PTM.TLBW  BTM...BW           -1.   BFE...BW            1.
    PTN.TNBW  BTN...BW           -1.   BXA...BW            1.
    PTR.TNBW  BTR...BW           -1.   BXH...BW            1.
    PTW.TNBW  BTW...BW           -1.   BFU...BW            1.
    PTZ.TNBW  BTZ...BW           -1.   BXU...BW            1.
    PTX.TSBW  BTX...BW           -1.   BFS...BW            1.
    PTY.TSBW  BTY...BW           -1.   BXS...BW            1.
    PTS.TTBW  BTS...BW           -1.   BPE...BW            1.
    PTT.TTBW  BTT...BW           -1.   BPT...BW            1.
    PTU.TUBW  BTU...BW           -1.   BPU...BW            1.
    PTV.TUBW  BTV...BW           -1.   BX6...BW            1.
    PSA.TXBW  BSA...BW           -1.   BXB...BW            1.
    PSB.TXBW  BSB...BW           -1.   BXL...BW            1.
    PSC.TXBW  BSC...BW           -1.   BFK...BW            1.
    PSD.TXBW  BSD...BW           -1.   BXK...BW            1.
    PPG.T1BW  BPG...BW           -1.   BZB...BW            1.
    PPH.T1BW  BPH...BW           -1.   BZC...BW            1.
    PTA.T1BW  BTA...BW           -1.   BZA...BW            1.
    PTD.T1BW  BTD...BW           -1.   BZD...BW            1.
    PTE.T1BW  BTE...BW           -1.   BZE...BW            1.
    PZA.T2BW  BZA...BW           -1.   BZF...BW            1.
    PZB.T2BW  BZB...BW           -1.   BZG...BW            1.
    PZC.T2BW  BZC...BW           -1.   BZH...BW            1.
    PZD.T2BW  BZD...BW           -1.   BZI...BW            1.
    PZE.T2BW  BZE...BW           -1.   BZJ...BW            1.
    PZF.2XBW  UFU...BW         -.011   UH2...BW        .98457
    PZF.2XBW  UNF...BW         -.007   U9R...BW         -.083
    PZF.2XBW  CTF.P2BW           -1.   CBF.2XBW           -1.
    PZF.2XBW  CTC.2XBW       -.00001   CTV.2XBW           -1.
    PZF.2XBW  BH....BW        .05008   BGS...BW        .07564
    PZF.2XBW  BLP...BW        .10016   BSI...BW        .05849
    PZF.2XBW  BSJ...BW        .07667   BZF...BW           -1.
    PZF.2XBW  BK1...BW          .674
    PZG.2XBW  UFU...BW         -.011   UH2...BW       1.04473
    PZG.2XBW  UNF...BW         -.007   U9R...BW         -.083
    PZG.2XBW  CTF.P2BW           -1.   CBF.2XBW           -1.
    PZG.2XBW  CTC.2XBW       -.00001   CTV.2XBW           -1.
    PZG.2XBW  BH....BW        .05314   BGS...BW        .05521
    PZG.2XBW  BLP...BW        .06964   BSI...BW        .03369
    PZG.2XBW  BSJ...BW        .04936   BZG...BW           -1.
    PZG.2XBW  BK2...BW         .7563
    PZH.2XBW  UFU...BW         -.011   UH2...BW       1.07383
    PZH.2XBW  UNF...BW         -.007   U9R...BW         -.083
    PZH.2XBW  CTF.P2BW           -1.   CBF.2XBW           -1.
    PZH.2XBW  CTC.2XBW       -.00001   CTV.2XBW           -1.
    PZH.2XBW  BH....BW        .05462   BGS...BW        .04572
    PZH.2XBW  BLP...BW        .05646   BSI...BW        .02506
    PZH.2XBW  BSJ...BW        .03991   BZH...BW           -1.
    PZH.2XBW  BK3...BW         .8099
    PZI.2XBW  UFU...BW         -.011   UH2...BW        .75848
    PZI.2XBW  UNF...BW         -.007   U9R...BW         -.083
    PZI.2XBW  CTF.P2BW           -1.   CBF.2XBW           -1.
    PZI.2XBW  CTC.2XBW           -1.   CTV.2XBW       -.00001
    PZI.2XBW  BH....BW        .03858   BGS...BW        .04265
    PZI.2XBW  BLP...BW        .05193   BSI...BW        .02378
    PZI.2XBW  BSJ...BW        .03703   BK4...BW          .825
    PZI.2XBW  BZI...BW           -1.
    PZJ.2XBW  UFU...BW         -.011   UH2...BW        .66746
    PZJ.2XBW  UNF...BW         -.007   U9R...BW         -.083
    PZJ.2XBW  CTF.P2BW           -1.   CBF.2XBW           -1.
    PZJ.2XBW  CTC.2XBW           -1.   CTV.2XBW       -.00001
    PZJ.2XBW  BH....BW        .03395   BGS...BW        .03823
    PZJ.2XBW  BLP...BW        .04606   BSI...BW          .021
    PZJ.2XBW  BSJ...BW        .03277   BK5...BW         .8455
    PZJ.2XBW  BZJ...BW           -1.
    QVO2G2BW  BG2...BW          .995   BG2.VOBW           -1.
    QVO2G2BW  B...ML..          .005   NGS.G2BW           -.5
    QVO2G2BW  NOI.G2BW        -78.12   NMO.G2BW          -50.
    QVO2G2BW  NRO.G2BW          -50.   N02.G2BW           -5.
    QVO2G2BW  XVP.G2BW         -13.8   NWU.G2BW         -104.
    QVO2G2BW  XCV.G2BW          -18.   X58.G2BW          -45.
    QVO2G2BW  X12.G2BW          -65.
    QWT2G2BW  BG2.WTBW           -1.   XSP.G2BW         -240.
    QVO2G4BW  BG4...BW          .995   BG4.VOBW           -1.
    QVO2G4BW  B...ML..          .005   NGS.G4BW           -.5
    QVO2G4BW  NOI.G4BW        -81.32   NMO.G4BW          -50.
    QVO2G4BW  NRO.G4BW          -50.   N02.G4BW           -5.
    QVO2G4BW  XVP.G4BW         -13.8   NWU.G4BW         -104.
    QVO2G4BW  XCV.G4BW          -18.   X58.G4BW          -45.
    QVO2G4BW  X12.G4BW          -65.
    QWT2G4BW  BG4.WTBW           -1.   XSP.G4BW         -240.
    QVO2G8BW  BG8...BW          .995   BG8.VOBW           -1.
    QVO2G8BW  B...ML..          .005   NGS.G8BW           -.5
    QVO2G8BW  NOI.G8BW        -82.59   NMO.G8BW          -50.
    QVO2G8BW  NRO.G8BW          -50.   N02.G8BW           -5.
    QVO2G8BW  XVP.G8BW         -13.8   NWU.G8BW         -104.
    QVO2G8BW  XCV.G8BW          -18.   X58.G8BW          -45.
    QVO2G8BW  X12.G8BW          -65.
    QWT2G8BW  BG8.WTBW           -1.   XSP.G8BW         -240.
    UH2...BW  BH....BW            1.
    U9R.BLBW  U9R...BW          .956
    U.S.BSBW  U.S...BW            1.
    URX.BXBW  URX...BW            1.
    UFU.BYBW  UFU...BW            1.
    UNF.BYBW  UNF...BW            1.
    U9R.GGBW  U9R...BW            1.   BGS...BW           -1.
    U9R.GHBW  U9R...BW          .744   BSN...BW           -1.
    U9R.GIBW  U9R...BW          .659   BSI...BW           -1.
    U9R.GLBW  U9R...BW          .696   BSR...BW           -1.
    U9R.GNBW  U9R...BW          .686   BSJ...BW           -1.
    U9R.GQBW  U9R...BW          .696   BSU...BW           -1.
    U9R.GUBW  U9R...BW         1.051   BYU...BW           -1.
    U9R.GVBW  U9R...BW          .708   BSQ...BW           -1.
    U9R.G2BW  U9R...BW            1.   BH....BW           -1.
    U9R.G2BW  UH2...BW        -19.66
    URX.SXBW  URX...BW           -1.
    WMO2G2BW  WMO.G2BW           -1.   NOI.G2BW           .52
    WMO2G2BW  NMO.G2BW            1.
    WRO2G2BW  WRO.G2BW           -1.   NOI.G2BW           .37
    WRO2G2BW  NRO.G2BW            1.
    WSU2G2BW  WSU.G2BW           -1.   XSP.G2BW        10000.
    WVP2G2BW  WVP.G2BW           -1.   XVP.G2BW            1.
    WVP2G2BW  XCV.G2BW            1.
    W022G2BW  W02.G2BW           -1.   N02.G2BW            1.
    W022G2BW  NWU.G2BW           .67
    W122G2BW  W12.G2BW           -1.   NWU.G2BW            1.
    W122G2BW  X12.G2BW            1.
    W582G2BW  W58.G2BW           -1.   XCV.G2BW           .13
    W582G2BW  X58.G2BW            1.
    WMO2G4BW  WMO.G4BW           -1.   NOI.G4BW           .56
    WMO2G4BW  NMO.G4BW            1.
    WRO2G4BW  WRO.G4BW           -1.   NOI.G4BW           .35
    WRO2G4BW  NRO.G4BW            1.
    WSU2G4BW  WSU.G4BW           -1.   XSP.G4BW        10000.
    WVP2G4BW  WVP.G4BW           -1.   XVP.G4BW            1.
    WVP2G4BW  XCV.G4BW            1.
    W022G4BW  W02.G4BW           -1.   N02.G4BW            1.
    W022G4BW  NWU.G4BW           .67
    W122G4BW  W12.G4BW           -1.   NWU.G4BW            1.
    W122G4BW  X12.G4BW            1.
    W582G4BW  W58.G4BW           -1.   XCV.G4BW           .13
    W582G4BW  X58.G4BW            1.
    WMO2G8BW  WMO.G8BW           -1.   NOI.G8BW           .59
    WMO2G8BW  NMO.G8BW            1.
    WRO2G8BW  WRO.G8BW           -1.   NOI.G8BW           .31
    WRO2G8BW  NRO.G8BW            1.
    WSU2G8BW  WSU.G8BW           -1.   XSP.G8BW        10000.
    WVP2G8BW  WVP.G8BW           -1.   XVP.G8BW            1.
    WVP2G8BW  XCV.G8BW            1.
    W022G8BW  W02.G8BW           -1.   N02.G8BW            1.
    W022G8BW  NWU.G8BW           .67
    W122G8BW  W12.G8BW           -1.   NWU.G8BW            1.
    W122G8BW  X12.G8BW            1.
    W582G8BW  W58.G8BW           -1.   XCV.G8BW           .13
    W582G8BW  X58.G8BW            1.
RHS
BOUNDS
 UP B1BOUND   CTP.A.BW            8.
 UP B1BOUND   CTF.ASBW           10.
 LO B1BOUND   CTF.C.BW          250.
 UP B1BOUND   CTF.C.BW          290.
 LO B1BOUND   CBF.F.BW           90.
 UP B1BOUND   CBF.F.BW          100.
 UP B1BOUND   CTF.FTBW           25.
 UP B1BOUND   CTF.F2BW          135.
 UP B1BOUND   CBF.P1BW           19.
 UP B1BOUND   CBF.P2BW           23.
 FX B1BOUND   CBF.2XBW            0.
 UP B1BOUND   DC3...BW           5.9
 UP B1BOUND   DC4...BW           5.1
 UP B1BOUND   DLP...BW          500.
 UP B1BOUND   DR1...BW          500.
 UP B1BOUND   DYU...BW          500.
 UP B1BOUND   D4....BW           16.
 UP B1BOUND   D4Z...BW          500.
 UP B1BOUND   D6....BW          500.
 UP B1BOUND   D8....BW          500.
 UP B1BOUND   D9....BW          500.
 UP B1BOUND   EBI...BW           50.
 UP B1BOUND   ECD...BW          330.
 UP B1BOUND   EHC...BW           15.
 FX B1BOUND   ELI...BW            0.
 UP B1BOUND   EPI...BW           35.
 UP B1BOUND   ETO...BW            2.
 UP B1BOUND   GBF.P2BW           10.
 UP B1BOUND   HBF.P1BW           10.
 FR B1BOUND   UH2...BW
 FR B1BOUND   U9R.BLBW
 FX B1BOUND   URX.BXBW       1100.41
 UP B1BOUND   U9R.GUBW            1.
ENDATA
