NAME          SCFXM1
ROWS
 N  .COSTA
 E  1DT001
 E  1DT002
 E  1DT003
 E  1DT004
 E  1DT005
 E  1DT006
 E  1DT007
 E  1DT008
 L  1DT009
 E  1DT010
 E  1DT011
 E  1DT012
 E  1DT013
 E  1DT014
 E  1DT015
 E  1DT016
 L  1DT017
 E  1DT018
 E  1DT021
 E  1DT022
 E  1DT023
 E  1DT024
 L  1DT025
 E  1DT026
 L  1DT027
 L  1DT028
 L  1DT029
 L  1DT030
 E  1DT031
 L  1DT032
 L  1DT033
 E  1DT034
 E  1DT035
 E  1DT036
 E  1DT037
 L  1DT038
 L  1DT039
 L  1DT040
 L  1DT041
 L  1DT042
 E  1DT043
 E  1DT044
 L  1DT045
 L  1DT047
 L  1DT048
 L  1DT049
 E  1DT050
 L  1DT051
 E  1DT052
 L  1DT053
 L  1DT055
 E  1DT056
 L  1DT057
 E  1DT059
 L  1DT060
 L  1DT062
 L  1DT063
 L  1DT064
 L  1DT066
 L  1DT067
 E  1DT068
 E  1DT069
 L  1DT070
 E  1DT071
 E  1DT072
 E  1DT073
 L  1DT074
 E  1DT075
 L  1DT076
 L  1DT077
 E  1DT078
 E  1DT079
 E  1DT080
 E  1DT081
 E  1DT082
 E  1DT083
 E  1DT084
 E  1DT087
 E  1DT088
 E  1DT089
 E  1DT090
 E  1DT091
 E  1DT092
 E  1DT094
 L  1DT095
 L  1DT096
 L  1DT097
 L  1DT098
 L  1DT099
 E  1DT100
 E  1DT101
 E  1DT102
 E  1DT019
 E  1DT020
 E  1DT046
 L  1DT093
 E  SCROW1
 E  SCROW2
 E  SCROW3
 E  SCROW4
 E  SCROW5
 E  1MS001
 E  1MS002
 E  1MS003
 E  1MS004
 E  1MS005
 E  1MS006
 E  1MS007
 E  1MS008
 E  1MS009
 E  1MS010
 E  1MS011
 E  1MS012
 E  1MS013
 E  1MS014
 E  1MS015
 E  1MS016
 E  1MS017
 E  1MS018
 E  1MS019
 E  1MS020
 E  1MS021
 L  1MS022
 E  1MS023
 E  1MS024
 E  1MS025
 E  1MS026
 E  1MS027
 E  1MS028
 E  1MS029
 E  1MS030
 E  1MS031
 E  1MS032
 E  1MS033
 E  1MS034
 E  1MS035
 E  1MS036
 L  1MS037
 E  1MS038
 E  1MS039
 E  1MS040
 E  1MS041
 L  1MS043
 L  1MS044
 L  1MS045
 L  1MS046
 L  1MS047
 L  1MS048
 L  1MS049
 L  1MS050
 L  1MS051
 L  1MS052
 L  1MS053
 L  1MS054
 E  1MS055
 E  1MS059
 L  1MS060
 L  1MS061
 L  1MS062
 L  1MS063
 L  1MS064
 E  1MS065
 E  1MS067
 E  1MS068
 E  1MS069
 E  1MS070
 E  1MS071
 E  1MS072
 E  1MS073
 E  1MS074
 L  1MS075
 L  1MS076
 L  1MS078
 L  1MS077
 E  1MS042
 E  1MS056
 E  1MS057
 E  1MS058
 E  SCROW6
 E  1DT054
 E  1DT058
 E  1DT061
 E  1DT065
 E  1PD004
 E  1PD009
 E  1PD010
 E  1PD011
 E  1PD012
 E  1PD013
 E  1PD014
 E  1PD015
 E  1PD016
 E  1PD017
 E  1PD018
 E  1PD019
 E  1PD020
 L  1PD021
 L  1PD022
 L  1PD023
 L  1PD024
 E  1PD025
 E  1PD026
 E  1PD027
 E  1PD028
 E  1PD029
 E  1PD030
 E  1PD031
 E  1PD032
 E  1PD033
 E  1PD034
 L  1PD035
 L  1PD036
 L  1PD037
 L  1PD038
 E  1PD039
 E  1PD040
 E  1PD049
 E  1PD050
 E  1PD051
 E  1PD052
 E  1PD053
 E  1PD054
 E  1PD055
 E  1PD056
 E  1PD057
 E  1PD058
 E  1PD059
 E  1PD060
 E  1PD061
 E  1PD062
 E  1PD063
 E  1PD064
 E  1PD065
 E  1PD066
 E  1PD067
 E  1PD068
 E  1PD069
 E  1PD070
 E  1PD071
 E  1PD072
 E  1RB045
 E  1RB038
 E  1RB039
 E  1RB040
 E  1RB041
 L  1RB001
 L  1RB002
 L  1RB003
 L  1RB004
 L  1RB005
 L  1RB006
 L  1RB059
 L  1RB060
 L  1RB061
 L  1RB063
 L  1RB064
 L  1RB065
 L  1RB066
 L  1RB067
 L  1RB068
 E  1RB070
 L  1RB073
 E  1RB076
 L  1RB077
 L  1RB079
 E  1RB080
 E  1RB083
 L  1RB084
 E  1RB086
 E  1RB087
 E  1RB088
 L  1RB092
 L  1RB007
 L  1RB008
 L  1RB009
 L  1RB010
 L  1RB011
 L  1RB012
 L  1RB013
 L  1RB014
 L  1RB015
 L  1RB016
 L  1RB017
 L  1RB018
 L  1RB019
 L  1RB020
 L  1RB021
 L  1RB022
 L  1RB023
 L  1RB024
 L  1RB025
 L  1RB026
 L  1RB027
 L  1RB028
 L  1RB029
 L  1RB030
 E  1RB031
 E  1RB032
 L  1RB033
 L  1RB034
 L  1RB035
 L  1RB036
 L  1RB037
 E  1RB042
 L  1RB043
 L  1RB044
 L  1RB046
 L  1RB047
 L  1RB048
 L  1RB049
 L  1RB050
 L  1RB051
 L  1RB052
 L  1RB053
 L  1RB054
 L  1RB055
 L  1RB056
 L  1RB057
 L  1RB058
 L  1RB062
 E  1RB069
 L  1RB071
 L  1RB072
 L  1RB078
 L  1RB081
 L  1RB082
 L  1RB085
 L  1RB089
 L  1RB090
 L  1RB091
COLUMNS
    1D1IK     1DT012              1.   1DT060            -.13
    1D1IK     SCROW4              1.
    1D1IN     SCROW4             -1.   1DT098              1.
    1D1MK     1DT011              1.   1DT060           -.154
    1D1MK     SCROW4              1.
    1D1YK     1DT010              1.   1DT060            .278
    1D1YK     SCROW4              1.
    1D2ID     1DT015              1.   1DT062           -.281
    1D2ID     1DT063             7.5   1DT064           -12.6
    1D2ID     SCROW5              1.
    1D2IK     1DT012              1.   1DT062          -.3333
    1D2IK     1DT063            -30.   1DT064           -6.25
    1D2IK     SCROW5              1.
    1D2IN     SCROW5             -1.   1DT099              1.
    1D2LC     1DT043              1.   1DT062            .013
    1D2LC     1DT063              5.   1DT064           10.25
    1D2LC     SCROW5              1.
    1D2MD     1DT014              1.   1DT062           -.316
    1D2MD     1DT063             20.   1DT064            -20.
    1D2MD     SCROW5              1.
    1D2MK     1DT011              1.   1DT062           -.354
    1D2MK     1DT063           -22.5   1DT064            -16.
    1D2MK     SCROW5              1.
    1D2PB     1DT059              1.   1DT062           -.394
    1D2PB     1DT063           -62.5   1DT064            -7.5
    1D2PB     SCROW5              1.
    1D2US     1DT026              1.   1DT062           -.307
    1D2US     1DT063           -17.5   1DT064             -2.
    1D2US     SCROW5              1.
    1D2YD     1DT013              1.   1DT062            .724
    1D2YD     1DT063             2.5   1DT064            -7.4
    1D2YD     SCROW5              1.
    1D2YK     1DT010              1.   1DT062            .073
    1D2YK     1DT063            -45.   1DT064             -4.
    1D2YK     SCROW5              1.
    1D4BY     1DT093             -1.   1DT101           -3.26
    1DAOU     1DT009          -.0137   1DT016          -.0012
    1DAOU     1DT021           .0232   1DT024              .1
    1DAOU     1DT025            -.83   1DT026            -.15
    1DAOU     1DT033              1.   1DT082           -.001
    1DAOU     1DT083              1.   1DT084          -.0051
    1DAOU     1DT088          -.0051   1DT089          -.0137
    1DAOU     1DT100          -.0836
    1DASP     1DT018            .138   1DT020            .862
    1DASP     1DT031              1.   1DT102            3.46
    1DBZS     1DT035              1.   1DT049              1.
    1DBZS     1DT102            8.95
    1DC2P     1DT001              1.   1DT021             -1.
    1DC3A     1DT046          -.7739   1DT047              1.
    1DC3A     1DT091           .4348   1DT092           .5652
    1DC3A     1DT100          -.3284
    1DC3P     1DT021            -.61   1DT090              1.
    1DC3R     1DT001            -.09   1DT002              1.
    1DC3R     1DT090            -.85
    1DC4A     1DT046             -.8   1DT047              1.
    1DC4A     1DT092           .5349   1DT094           .4651
    1DC4A     1DT100          -.3399
    1DC65     .COSTA             10.   1DT001           -.025
    1DC65     1DT002           -.036   1DT003           -.058
    1DC65     1DT004           -.041   1DT005           -.011
    1DC65     1DT006           -.077   1DT016           -.475
    1DC65     1DT021           .0318   1DT025              1.
    1DC65     1DT042            -.05   1DT043            -.27
    1DC65     1DT044            -.08   1DT045              1.
    1DC75     1DT001           -.036   1DT002           -.042
    1DC75     1DT003           -.067   1DT004           -.047
    1DC75     1DT005           -.015   1DT006            -.09
    1DC75     1DT016           -.525   1DT021           .0318
    1DC75     1DT025              1.   1DT042          -.0525
    1DC75     1DT043            -.22   1DT044            -.03
    1DC75     1DT045              1.   1DT100          -.0124
    1DC8U     1DT084             -1.   1DT089              1.
    1DCKP     1DT017              1.   1DT021          -4.409
    1DCKP     1DT022              1.
    1DCKS     1DT022              1.   1DT102              4.
    1DCOU     1DT021           .0232   1DT024             .55
    1DCOU     1DT025             -.9   1DT026             -.1
    1DCOU     1DT033              1.   1DT081              1.
    1DCOU     1DT082          -.0039   1DT100          -.0836
    1DCTL     1DT016              1.   1DT042              1.
    1DCTL     1DT043             -1.
    1DDID     1DT015              1.   1DT066           -.282
    1DDID     1DT067              1.   1DT068              1.
    1DDMD     1DT014              1.   1DT066           -.316
    1DDMD     1DT067              1.   1DT068              1.
    1DDUS     1DT026              1.   1DT066           -.307
    1DDUS     1DT067             -3.   1DT068              1.
    1DDYD     1DT013              1.   1DT066            .724
    1DDYD     1DT067              1.   1DT068              1.
    1DFVB     1DT001           -.017   1DT002           -.016
    1DFVB     1DT003            -.01   1DT004           -.003
    1DFVB     1DT005           -.007   1DT006           -.007
    1DFVB     1DT008            -.15   1DT009          -.1074
    1DFVB     1DT016          -.0079   1DT019              1.
    1DFVB     1DT021            .038   1DT072            .676
    1DFVB     1DT074              1.   1DT084          -.0427
    1DFVB     1DT088          -.0427   1DT089          -.1074
    1DFVB     1DT102            2.21
    1DGBY     1DT025             -1.   1DT033              1.
    1DGBY     1DT077              1.   1DT101           -2.88
    1DGOU     1DT009          -.0137   1DT016          -.0012
    1DGOU     1DT018              1.   1DT021           .0232
    1DGOU     1DT024              .2   1DT025            -.88
    1DGOU     1DT026             -.1   1DT033              1.
    1DGOU     1DT082          -.0039   1DT084          -.0051
    1DGOU     1DT088          -.0051   1DT089          -.0137
    1DGOU     1DT100          -.0836
    1DH2P     1DT021          -.0511   1DT024              1.
    1DHNU     1DT009           -.785   1DT021           .0232
    1DHNU     1DT024             .45   1DT033              1.
    1DHNU     1DT080              1.   1DT084           -.215
    1DHNU     1DT088           -.215   1DT089           -.785
    1DHNU     1DT100          -.0836
    1DHPU     1DT021           .0244   1DT024             .05
    1DHPU     1DT041              1.   1DT084              1.
    1DHPU     1DT100          -.0409
    1DI4R     1DT001          -.0522   1DT004              1.
    1DI4R     1DT092            -.92
    1DIBY     1DT092             -1.   1DT101           -3.67
    1DICD     .COSTA             10.   1DT002           -.007
    1DICD     1DT004          -.0034   1DT005          -.0145
    1DICD     1DT007          -.0195   1DT009           -.098
    1DICD     1DT012           -.179   1DT015            -.08
    1DICD     1DT019            -.15   1DT021           .0139
    1DICD     1DT025           -.215   1DT028              1.
    1DICD     1DT078           -.223   1DT083             -.1
    1DICD     1DT088           -.125
    1DICG     1DT002           -.007   1DT004          -.0034
    1DICG     1DT005          -.0145   1DT007          -.0195
    1DICG     1DT009           -.128   1DT012           -.149
    1DICG     1DT015           -.045   1DT019            -.15
    1DICG     1DT021           .0139   1DT025            -.26
    1DICG     1DT028              1.   1DT078           -.253
    1DICG     1DT083            -.09   1DT088           -.125
    1DICG     1DT100          -.0487   1DT101          -3.114
    1DICK     .COSTA             10.   1DT002           -.007
    1DICK     1DT004          -.0034   1DT005          -.0145
    1DICK     1DT007          -.0195   1DT009           -.098
    1DICK     1DT012           -.179   1DT015           -.045
    1DICK     1DT019            -.15   1DT021           .0139
    1DICK     1DT025            -.26   1DT028              1.
    1DICK     1DT078           -.223   1DT083            -.09
    1DICK     1DT088           -.125
    1DIGD     .COSTA             10.   1DT002           -.007
    1DIGD     1DT004          -.0034   1DT005          -.0145
    1DIGD     1DT007          -.0195   1DT009           -.128
    1DIGD     1DT012           -.149   1DT015            -.08
    1DIGD     1DT019            -.15   1DT021           .0139
    1DIGD     1DT025           -.215   1DT028              1.
    1DIGD     1DT078           -.253   1DT083             -.1
    1DIGD     1DT088           -.125
    1DLBY     1DT002            -.02   1DT004            -.38
    1DLBY     1DT005             -.3   1DT007             -.2
    1DLBY     1DT076              1.   1DT101           -2.31
    1DLPU     1DT021           .0244   1DT024             .05
    1DLPU     1DT040              1.   1DT078              1.
    1DLPU     1DT100          -.0271
    1DMCD     .COSTA             10.   1DT002          -.0072
    1DMCD     1DT004          -.0068   1DT005          -.0375
    1DMCD     1DT007           -.023   1DT009            -.09
    1DMCD     1DT011           -.202   1DT014           -.095
    1DMCD     1DT019           -.115   1DT021           .0139
    1DMCD     1DT025          -.2285   1DT028              1.
    1DMCD     1DT030              1.   1DT078           -.205
    1DMCD     1DT083            -.08   1DT087           -.115
    1DMCG     1DT002          -.0072   1DT004          -.0068
    1DMCG     1DT005          -.0375   1DT007           -.023
    1DMCG     1DT009            -.13   1DT011           -.162
    1DMCG     1DT014            -.06   1DT019           -.115
    1DMCG     1DT021           .0139   1DT025          -.2735
    1DMCG     1DT028              1.   1DT030              1.
    1DMCG     1DT078           -.245   1DT083            -.07
    1DMCG     1DT087           -.115   1DT100          -.0487
    1DMCG     1DT101         -2.9375
    1DMCK     .COSTA             10.   1DT002          -.0072
    1DMCK     1DT004          -.0068   1DT005          -.0375
    1DMCK     1DT007           -.023   1DT009            -.09
    1DMCK     1DT011           -.202   1DT014            -.06
    1DMCK     1DT019           -.115   1DT021           .0139
    1DMCK     1DT025          -.2735   1DT028              1.
    1DMCK     1DT030              1.   1DT078           -.205
    1DMCK     1DT083            -.07   1DT087           -.115
    1DMCST    .COSTA              1.   1DT101              1.
    1DMGD     .COSTA             10.   1DT002          -.0072
    1DMGD     1DT004          -.0068   1DT005          -.0375
    1DMGD     1DT007           -.023   1DT009            -.13
    1DMGD     1DT011           -.162   1DT014           -.095
    1DMGD     1DT019           -.115   1DT021           .0139
    1DMGD     1DT025          -.2285   1DT028              1.
    1DMGD     1DT030              1.   1DT078           -.245
    1DMGD     1DT083            -.08   1DT087           -.115
    1DN4R     1DT001          -.0472   1DT005              1.
    1DN4R     1DT093            -.93
    1DNBY     1DT021             -1.   1DT101           -2.67
    1DO3R     1DT001           -.118   1DT003              1.
    1DO3R     1DT091            -.82
    1DO4N     1DT093             -1.   1DT094              1.
    1DO4R     1DT001          -.0138   1DT006              1.
    1DO4R     1DT094            -.98
    1DOCST    .COSTA              1.   1DT100              1.
    1DP94     1DT001          -.0637   1DT002         -.03806
    1DP94     1DT004         -.02075   1DT005         -.00309
    1DP94     1DT009            1.15   1DT021           .0635
    1DP94     1DT024          -.7938   1DT038            1.15
    1DP94     1DT051             4.3   1DT052            -3.6
    1DP94     1DT053             -2.   SCROW2              1.
    1DP94     1DT059         -.06902   1DT100          -.0735
    1DP94     1DT101          -.3717
    1DP98     1DT001         -.09265   1DT002         -.04946
    1DP98     1DT004         -.02731   1DT005         -.00303
    1DP98     1DT009           1.207   1DT021           .0635
    1DP98     1DT024          -.9296   1DT038           1.412
    1DP98     1DT051             1.9   1DT052             -2.
    1DP98     1DT053             -2.   SCROW2              1.
    1DP98     1DT059         -.07243   1DT100          -.0771
    1DP98     1DT101          -.3717
    1DPAL     1DT046              1.   1DT051            -8.4
    1DPAL     1DT052            3.82   1DT053             -2.
    1DPAL     SCROW2              1.   1DT101          -.3717
    1DPC4     1DT051            -8.3   1DT052           38.43
    1DPC4     1DT053             -2.   SCROW2              1.
    1DPC4     1DT093              1.   1DT101          -.3717
    1DPCT     1DT016              1.   1DT051             5.5
    1DPCT     1DT052           -3.31   1DT053             -2.
    1DPCT     SCROW2              1.   1DT101          -.3717
    1DPIN     SCROW2             -1.   1DT096              1.
    1DPOH     1DT034              1.   1DT051             -.4
    1DPOH     1DT052           10.15   1DT053             -2.
    1DPOH     SCROW2              1.   1DT101          -.3717
    1DPPS     1DT048              1.   1DT090             .95
    1DPPS     1DT091             .05   1DT102            2.41
    1DPREV    .COSTA             -1.   1DT102             -1.
    1DPRF     1DT007              1.   1DT051             6.7
    1DPRF     1DT052             .26   1DT053             -2.
    1DPRF     SCROW2              1.   1DT101          -.3717
    1DPTL     1DT051            1.25   1DT053              1.
    1DPTL     1DT101          .09293
    1DPXY     1DT037              1.   1DT051           -20.3
    1DPXY     1DT052            -8.5   1DT053             -2.
    1DPXY     SCROW2              1.   1DT101          -.3717
    1DPYS     1DT027              1.   1DT090             .02
    1DPYS     1DT091             .98   1DT102            4.72
    1DR94     1DT001          -.0637   1DT002         -.03806
    1DR94     1DT004         -.02075   1DT005         -.00309
    1DR94     1DT009            1.15   1DT021           .0635
    1DR94     1DT024          -.7938   1DT038            1.15
    1DR94     1DT055             -.4   1DT056            -3.6
    1DR94     1DT057            -1.5   SCROW3              1.
    1DR94     1DT059         -.06902   1DT100          -.0735
    1DR94     1DT101          -.3717
    1DR98     1DT001         -.09265   1DT002         -.04946
    1DR98     1DT004         -.02731   1DT005         -.00303
    1DR98     1DT009           1.207   1DT021           .0635
    1DR98     1DT024          -.9296   1DT038           1.412
    1DR98     1DT055            -2.5   1DT056             -2.
    1DR98     1DT057            -1.5   SCROW3              1.
    1DR98     1DT059         -.07243   1DT100          -.0771
    1DR98     1DT101          -.3717
    1DRAL     1DT046              1.   1DT055            -9.4
    1DRAL     1DT056            3.82   1DT057            -1.5
    1DRAL     SCROW3              1.   1DT101          -.3717
    1DRBT     1DT001           -.021   1DT002           -.025
    1DRBT     1DT004           -.013   1DT005           -.026
    1DRBT     1DT007           -.451   1DT021            .115
    1DRBT     1DT024             -.9   1DT034           -.095
    1DRBT     1DT035           -.098   1DT036           -.182
    1DRBT     1DT037            -.06   1DT039              1.
    1DRBT     1DT078             -1.   1DT100          -.0639
    1DRBT     SCROW1             -1.
    1DRC4     1DT055            -9.9   1DT056           38.43
    1DRC4     1DT057            -1.5   SCROW3              1.
    1DRC4     1DT093              1.   1DT101          -.3717
    1DRCT     1DT016              1.   1DT055            -2.6
    1DRCT     1DT056           -3.31   1DT057            -1.5
    1DRCT     SCROW3              1.   1DT101          -.3717
    1DRIN     SCROW3             -1.   1DT097              1.
    1DROH     1DT034              1.   1DT055            -1.7
    1DROH     1DT056           10.15   1DT057            -1.5
    1DROH     SCROW3              1.   1DT101          -.3717
    1DRRF     1DT007              1.   1DT055             5.6
    1DRRF     1DT056             .26   1DT057            -1.5
    1DRRF     SCROW3              1.   1DT101          -.3717
    1DRTL     1DT055            1.87   1DT057              1.
    1DRTL     1DT101          .09293
    1DRXY     1DT037              1.   1DT055           -24.1
    1DRXY     1DT056            -8.5   1DT057            -1.5
    1DRXY     SCROW3              1.   1DT101          -.3717
    1DS5B     1DT019              1.   1DT073            2.64
    1DS5B     1DT075              1.
    1DS5F     1DT020              1.   1DT073            5.69
    1DS5F     1DT075              1.
    1DS5G     1DT023              1.   1DT073            6.49
    1DS5G     1DT075              1.
    1DS5L     1DT043              1.   1DT073           -7.69
    1DS5L     1DT075              1.
    1DS5P     1DT021             -1.   1DT075             -1.
    1DS5R     1DT072             -1.   1DT073              .3
    1DS5R     1DT075              1.   1DT102           -3.36
    1DS5S     1DT044              1.   1DT073           -3.69
    1DS5S     1DT075              1.
    1DS5V     1DT008              1.   1DT073           -8.08
    1DS5V     1DT075              1.
    1DS6B     1DT019              1.   1DT069            3.62
    1DS6B     1DT070           -.487   1DT071              1.
    1DS6L     1DT043              1.   1DT069           -6.69
    1DS6L     1DT070           -.428   1DT071              1.
    1DS6R     1DT069            1.28   1DT070           -.151
    1DS6R     1DT071              1.   1DT072             -1.
    1DS6R     1DT102           -3.36
    1DS6S     1DT044              1.   1DT069            -2.7
    1DS6S     1DT070          -.0946   1DT071              1.
    1DS6V     1DT008              1.   1DT069           -7.09
    1DS6V     1DT070           -5.04   1DT071              1.
    1DSCO     1DT001           -.072   1DT002           -.027
    1DSCO     1DT003           -.034   1DT004           -.005
    1DSCO     1DT005           -.015   1DT006           -.019
    1DSCO     1DT009           -.005   1DT016            -.04
    1DSCO     1DT019              1.   1DT021           .0475
    1DSCO     1DT022           -.012   1DT023            -.04
    1DSCO     1DT032              1.   1DT080           -.137
    1DSCO     1DT081            -.58   1DT084           -.055
    1DSCO     1DT088           -.055   1DT089           -.005
    1DSCO     1DT100          -.1415
    1DSPF     1DT021          -1.055   1DT023              1.
    1DSUS     1DT082              1.   1DT102             25.
    1DSVB     1DT001           -.035   1DT002           -.026
    1DSVB     1DT003           -.016   1DT004           -.004
    1DSVB     1DT005            -.01   1DT006           -.001
    1DSVB     1DT008             -.1   1DT009          -.1428
    1DSVB     1DT016          -.0105   1DT021            .038
    1DSVB     1DT044              1.   1DT072             .61
    1DSVB     1DT074              1.   1DT084          -.0567
    1DSVB     1DT088          -.0567   1DT089          -.1428
    1DSVB     1DT102           1.977
    1DTLS     1DT036              1.   1DT049              1.
    1DTLS     1DT102            6.78
    1DV6I     1DT001           -.021   1DT002           -.025
    1DV6I     1DT004           -.013   1DT005           -.026
    1DV6I     1DT007           -.451   1DT021            .115
    1DV6I     1DT024             -.9   1DT034           -.095
    1DV6I     1DT035           -.098   1DT036           -.182
    1DV6I     1DT037            -.06   1DT039              1.
    1DV6I     1DT088              1.   1DT100          -.0639
    1DV6M     1DT001           -.029   1DT002            -.03
    1DV6M     1DT004           -.022   1DT005           -.033
    1DV6M     1DT007            -.52   1DT021            .115
    1DV6M     1DT024            -.55   1DT034           -.133
    1DV6M     1DT035           -.042   1DT036           -.125
    1DV6M     1DT037            -.01   1DT039              1.
    1DV6M     1DT087              1.   1DT100          -.0639
    1DV6Y     1DT001           -.025   1DT002           -.027
    1DV6Y     1DT004           -.015   1DT005           -.024
    1DV6Y     1DT007            -.48   1DT021            .115
    1DV6Y     1DT024            -.75   1DT034           -.124
    1DV6Y     1DT035            -.05   1DT036            -.15
    1DV6Y     1DT037            -.04   1DT039              1.
    1DV6Y     1DT079              1.   1DT100          -.0639
    1DXC3     1DT090              1.   1DT095              1.
    1DXI4     1DT092              1.
    1DXO4     1DT094              1.
    1DXPY     1DT091              1.
    1DXYS     1DT037              1.   1DT050              1.
    1DXYS     1DT102            7.98
    1DYCD     .COSTA             10.   1DT002           -.002
    1DYCD     1DT004          -.0025   1DT005          -.0093
    1DYCD     1DT007           -.005   1DT009            -.06
    1DYCD     1DT010           -.115   1DT013           -.075
    1DYCD     1DT018           -.431   1DT020            -.25
    1DYCD     1DT021           .0139   1DT028              1.
    1DYCD     1DT029              1.   1DT078            -.11
    1DYCD     1DT079            -.05
    1DYCG     1DT002          -.0022   1DT004          -.0025
    1DYCG     1DT005          -.0093   1DT007           -.005
    1DYCG     1DT009           -.075   1DT010             -.1
    1DYCG     1DT013            -.05   1DT018           -.456
    1DYCG     1DT020            -.25   1DT021           .0139
    1DYCG     1DT028              1.   1DT029              1.
    1DYCG     1DT078           -.125   1DT079            -.05
    1DYCG     1DT100          -.0487   1DT101         -2.6498
    1DYCK     .COSTA             10.   1DT002           -.002
    1DYCK     1DT004          -.0025   1DT005          -.0093
    1DYCK     1DT007           -.005   1DT009            -.06
    1DYCK     1DT010           -.115   1DT013            -.05
    1DYCK     1DT018           -.456   1DT020            -.25
    1DYCK     1DT021           .0139   1DT028              1.
    1DYCK     1DT029              1.   1DT078            -.11
    1DYCK     1DT079            -.05
    1DYCO     1DT001           -.135   1DT002            -.04
    1DYCO     1DT003            -.06   1DT004           -.011
    1DYCO     1DT005           -.022   1DT006           -.022
    1DYCO     1DT009           -.005   1DT016            -.04
    1DYCO     1DT020              1.   1DT021           .0475
    1DYCO     1DT022          -.0434   1DT023            -.04
    1DYCO     1DT032              1.   1DT080           -.136
    1DYCO     1DT081           -.307   1DT084           -.055
    1DYCO     1DT088           -.055   1DT089           -.005
    1DYCO     1DT100          -.1415
    1DYGD     .COSTA             10.   1DT002           -.002
    1DYGD     1DT004          -.0025   1DT005          -.0093
    1DYGD     1DT007           -.005   1DT009           -.075
    1DYGD     1DT010             -.1   1DT013           -.075
    1DYGD     1DT018           -.431   1DT020            -.25
    1DYGD     1DT021           .0139   1DT028              1.
    1DYGD     1DT029              1.   1DT078           -.125
    1DYGD     1DT079            -.05
    SCCOL1    SCROW1              1.   SCROW6              1.
    SCCOL2    SCROW2             -1.   1DT054              1.
    SCCOL3    SCROW3             -1.   1DT058              1.
    SCCOL4    SCROW4             -1.   1DT061              1.
    SCCOL5    SCROW5             -1.   1DT065              1.
    1M1SK     1MS014              1.   1MS056              1.
    1M1YK     1MS013              1.   1MS056              1.
    1M2BD     1MS027              1.   1MS046              2.
    1M2BD     1MS047            -40.   1MS048             -.4
    1M2BD     1MS057              1.
    1M2BG     1MS026              1.   1MS046             -1.
    1M2BG     1MS047            -40.   1MS048           -.403
    1M2BG     1MS057              1.
    1M2SK     1MS014              1.   1MS046            -15.
    1M2SK     1MS047            -35.   1MS048           -.288
    1M2SK     1MS057              1.
    1M2YK     1MS013              1.   1MS046            -10.
    1M2YK     1MS047            -40.   1MS048            .125
    1M2YK     1MS057              1.
    1M1IN     1MS056             -1.   1MS077              1.
    1M2IN     1MS057             -1.   1MS078              1.
    1M2LS     1MS033              1.   1MS046            17.5
    1M2LS     1MS047             5.5   1MS048           -.091
    1M2LS     1MS057              1.
    1M2LY     1MS032              1.   1MS046             15.
    1M2LY     1MS048           1.191   1MS057              1.
    1M2SD     1MS018              1.   1MS046            -16.
    1M2SD     1MS047             30.   1MS048           -.262
    1M2SD     1MS057              1.
    1M2SG     1MS016              1.   1MS046           -14.3
    1M2SG     1MS047             50.   1MS048           -.261
    1M2SG     1MS057              1.
    1M2UD     1MS046             -9.   1MS047              5.
    1M2UD     1MS048            -.26   1MS057              1.
    1M2UD     1MS059              1.
    1M2UG     1MS035              1.   1MS046            -9.5
    1M2UG     1MS047             20.   1MS048            -.21
    1M2UG     1MS057              1.
    1M2UL     1MS036              1.   1MS046            12.5
    1M2UL     1MS047             12.   1MS048            -.25
    1M2UL     1MS057              1.
    1M2YD     1MS017              1.   1MS046             -8.
    1M2YD     1MS047             20.   1MS048           1.064
    1M2YD     1MS057              1.
    1M2YG     1MS015              1.   1MS046            -8.5
    1M2YG     1MS047             30.   1MS048           1.115
    1M2YG     1MS057              1.
    1M5LS     1MS033              1.   1MS045           -7.11
    1M5LS     1MS055             -1.
    1M5LY     1MS032              1.   1MS045           -6.33
    1M5LY     1MS055             -1.
    1M5R      1MS041              1.   1MS055              1.
    1M5SB     1MS022              1.   1MS045            2.72
    1M5SB     1MS055             -1.
    1M5SD     1MS018              1.   1MS045           -4.13
    1M5SD     1MS055             -1.
    1M5SG     1MS016              1.   1MS045           -4.03
    1M5SG     1MS055             -1.
    1M5SL     1MS034              1.   1MS045           -3.58
    1M5SL     1MS055             -1.
    1M5YB     1MS021              1.   1MS045            5.17
    1M5YB     1MS055             -1.
    1M5YD     1MS017              1.   1MS045           -3.73
    1M5YD     1MS055             -1.
    1M5YG     1MS015              1.   1MS045           -3.83
    1M5YG     1MS055             -1.
    1M65S     1MS002           -.065   1MS004           -.025
    1M65S     1MS005           -.036   1MS006           -.056
    1M65S     1MS007          -.0175   1MS008           -.048
    1M65S     1MS009           -.064   1MS020              1.
    1M65S     1MS023            .027   1MS031            -.47
    1M65S     1MS033            -.27   1MS034            -.08
    1M65S     1MS063              1.
    1M65Y     1MS002           -.065   1MS004          -.0275
    1M65Y     1MS005           -.035   1MS006           -.056
    1M65Y     1MS007          -.0165   1MS008          -.0485
    1M65Y     1MS009           -.062   1MS019              1.
    1M65Y     1MS023            .027   1MS031            -.47
    1M65Y     1MS032            -.23   1MS034            -.12
    1M65Y     1MS063              1.
    1M6SB     1MS003            3.04   1MS022              1.
    1M6SB     1MS040              1.   1MS043            3.03
    1M6SB     1MS044            .151
    1M6SD     1MS003            3.04   1MS018              1.
    1M6SD     1MS040              1.   1MS043           -3.82
    1M6SD     1MS044           -.895
    1M6SG     1MS003            3.04   1MS016              1.
    1M6SG     1MS040              1.   1MS043           -3.72
    1M6SG     1MS044           -.891
    1M6SL     1MS003            3.04   1MS034              1.
    1M6SL     1MS040              1.   1MS043           -3.27
    1M6SL     1MS044            -.74
    1M75S     1MS002           -.065   1MS004           -.036
    1M75S     1MS005           -.041   1MS006          -.0665
    1M75S     1MS007            -.02   1MS008          -.0555
    1M75S     1MS009          -.0755   1MS020              1.
    1M75S     1MS023            .035   1MS031            -.52
    1M75S     1MS033            -.21   1MS034            -.04
    1M75S     1MS063              1.
    1M75Y     1MS002           -.065   1MS004           -.039
    1M75Y     1MS005           -.039   1MS006           -.064
    1M75Y     1MS007          -.0185   1MS008           -.054
    1M75Y     1MS009            -.07   1MS019              1.
    1M75Y     1MS023            .035   1MS031            -.52
    1M75Y     1MS032            -.18   1MS034            -.07
    1M75Y     1MS063              1.
    1M94S     1MS002            -.08   1MS004          -.0074
    1M94S     1MS005          -.0275   1MS007         -.00245
    1M94S     1MS008           -.014   1MS023           .0663
    1M94S     1MS025              1.   1MS028           -.885
    1M94S     1MS029          -.8785   1MS062              1.
    1M94Y     1MS002            -.11   1MS004          -.0128
    1M94Y     1MS005          -.0457   1MS007         -.00295
    1M94Y     1MS008         -.00231   1MS023           .0663
    1M94Y     1MS024              1.   1MS028             -.7
    1M94Y     1MS029          -.8442   1MS062              1.
    1M98S     1MS002            -.09   1MS004          -.0083
    1M98S     1MS005          -.0314   1MS007          -.0024
    1M98S     1MS008          -.0163   1MS023            .069
    1M98S     1MS025              1.   1MS028           -.925
    1M98S     1MS030          -.8606   1MS062            1.17
    1M98Y     1MS002            -.12   1MS004          -.0137
    1M98Y     1MS005            -.05   1MS007          -.0028
    1M98Y     1MS008          -.0255   1MS023            .069
    1M98Y     1MS024              1.   1MS028            -.74
    1M98Y     1MS030           -.824   1MS062            1.17
    1MAF1     1MS014             .75   1MS026             .25
    1MAF1     1MS039              1.
    1MAF2     1MS014             .75   1MS027             .25
    1MAF2     1MS039              1.
    1MAF3     1MS013             .75   1MS026             .25
    1MAF3     1MS039              1.
    1MAF4     1MS013             .75   1MS027             .25
    1MAF4     1MS039              1.
    1MALK     1DT046              1.   1MS001          -.3928
    1MALK     1MS042              1.   1MS049            -2.1
    1MALK     1MS050           -5.56   1MS051             -2.
    1MALK     1MS074            -.12
    1MASP     1MS003            3.36   1MS019            .138
    1MASP     1MS021            .862   1MS038              1.
    1MC2P     1MS004              1.   1MS065             -1.
    1MC3F     1MS023            -.67   1MS068              1.
    1MC3P     1MS005              1.   1MS065            -.09
    1MC3P     1MS068            -.85
    1MC3S     1MS003            2.31   1MS037              1.
    1MC3S     1MS068              1.
    1MC4P     1MS007              1.   1MS065          -.0472
    1MC4P     1MS067            -.93
    1MCSD     1MS001          -3.212   1MS002          -.0339
    1MCSD     1MS005           -.002   1MS007           -.007
    1MCSD     1MS008           -.004   1MS010           -.083
    1MCSD     1MS014            -.13   1MS018             -.1
    1MCSD     1MS020           -.344   1MS022            -.11
    1MCSD     1MS023           .0139   1MS061              1.
    1MCSD     1MS064             .13   1MS073            -.22
    1MCSG     1MS001          -3.212   1MS002          -.0339
    1MCSG     1MS005           -.002   1MS007           -.007
    1MCSG     1MS008           -.004   1MS010           -.083
    1MCSG     1MS012            -.22   1MS014            -.13
    1MCSG     1MS016            -.05   1MS020           -.394
    1MCSG     1MS022            -.11   1MS023           .0139
    1MCSG     1MS061              1.   1MS064             .13
    1MCYD     1MS001         -2.7289   1MS002          -.0493
    1MCYD     1MS005           -.001   1MS007           -.006
    1MCYD     1MS008           -.002   1MS010            -.03
    1MCYD     1MS013           -.107   1MS017             -.1
    1MCYD     1MS019           -.354   1MS021           -.275
    1MCYD     1MS023           .0139   1MS060              1.
    1MCYD     1MS061              1.   1MS064            .107
    1MCYD     1MS072           -.125
    1MCYG     1MS001         -2.7289   1MS002          -.0493
    1MCYG     1MS005           -.001   1MS007           -.006
    1MCYG     1MS008           -.002   1MS010            -.03
    1MCYG     1MS011           -.125   1MS013           -.107
    1MCYG     1MS015           -.047   1MS019           -.407
    1MCYG     1MS021           -.275   1MS023           .0139
    1MCYG     1MS060              1.   1MS061              1.
    1MCYG     1MS064            .107
    1MFG      1MS023             -1.   1MS065              1.
    1MH2P     1MS028              1.   1MS065          -.0511
    1MI4D     1DT093             -1.   1MS069              1.
    1MI4D     1MS074            -.12
    1MI4P     1MS008              1.   1MS065          -.0522
    1MI4P     1MS069            -.92
    1MIN4     1MS067             -1.   1MS069              1.
    1MMLC     .COSTA              1.   1MS001              1.
    1MN4D     1DT093              1.   1MS007             -1.
    1MN4D     1MS074            -.12
    1MO4P     1MS009              1.   1MS065          -.0138
    1MO4P     1MS071            -.98
    1MON4     1MS067             -1.   1MS071              1.
    1MOPC     .COSTA              1.   1MS002              1.
    1MP94     1MS001          -.3928   1MS029              1.
    1MP94     1MS042              1.   1MS049            -4.5
    1MP94     1MS050             2.2   1MS051             -2.
    1MP98     1MS001          -.3928   1MS030              1.
    1MP98     1MS042              1.   1MS049            -1.7
    1MP98     1MS050             -.5   1MS051             -2.
    1MPC3     1MS002            -.48   1MS005             -.1
    1MPC3     1MS031           -.625   1MS070              1.
    1MPC4     1MS002           -.408   1MS007             -.1
    1MPC4     1MS031           -.735   1MS071              1.
    1MPCP     1MS001          -.3928   1MS031              1.
    1MPCP     1MS042              1.   1MS049            -1.7
    1MPCP     1MS050             2.8   1MS051             -2.
    1MPIN     1MS042             -1.   1MS075              1.
    1MPN4     1MS001          -.3928   1MS042              1.
    1MPN4     1MS049            60.5   1MS050           -9.19
    1MPN4     1MS051             -2.   1MS067              1.
    1MPRC     .COSTA             -1.   1MS003             -1.
    1MPSD     1MS010            -.03   1MS025            -.75
    1MPSD     1MS027            -.22   1MS073              1.
    1MPSG     1MS010            -.03   1MS012              1.
    1MPSG     1MS025            -.86   1MS026            -.11
    1MPSR     1MS001          -.3928   1MS010              1.
    1MPSR     1MS042              1.   1MS049            -1.9
    1MPSR     1MS050             4.3   1MS051             -2.
    1MPTL     1MS001          .09821   1MS050             .87
    1MPTL     1MS051              1.
    1MPYD     1MS010            -.03   1MS024            -.81
    1MPYD     1MS027            -.16   1MS072              1.
    1MPYG     1MS010            -.03   1MS011              1.
    1MPYG     1MS024            -.89   1MS026            -.08
    1MPYP     1MS006              1.   1MS065           -.118
    1MPYP     1MS070            -.82
    1MR94     1MS029              1.   1MS052            -4.5
    1MR94     1MS053            -1.5   1MS054            -1.5
    1MR94     1MS058              1.
    1MR98     1MS030              1.   1MS052            -1.7
    1MR98     1MS053            -4.8   1MS054            -1.5
    1MR98     1MS058              1.
    1MRCP     1MS031              1.   1MS052            -1.7
    1MRCP     1MS053            -3.5   1MS054            -1.5
    1MRCP     1MS058              1.
    1MRIN     1MS001          -.3928   1MS058             -1.
    1MRIN     1MS076              1.
    1MRN4     1MS052            60.5   1MS053          -11.93
    1MRN4     1MS054            -1.5   1MS058              1.
    1MRN4     1MS067              1.
    1MRSR     1MS010              1.   1MS052            -1.9
    1MRSR     1MS053            3.02   1MS054            -1.5
    1MRSR     1MS058              1.
    1MRTL     1MS001          .09821   1MS053            1.43
    1MS6P     1MS023             -1.   1MS055              1.
    1MSRD     1DT020             -1.   1MS021              1.
    1MSRD     1MS074            -.12
    1MSWD     1DT019             -1.   1MS022              1.
    1MSWD     1MS074            -.12
    1MTCS     .COSTA             -1.   1MS074              1.
    1MULS     1MS002            -.04   1MS010            -.01
    1MULS     1MS023            .025   1MS028              .4
    1MULS     1MS033              1.   1MS036            -.99
    1MULS     1MS064              1.
    1MULY     1MS002            -.04   1MS010            -.01
    1MULY     1MS023            .025   1MS028              .5
    1MULY     1MS032              1.   1MS036            -.99
    1MULY     1MS064              1.
    1MUYD     1MS002            -.04   1MS010            -.01
    1MUYD     1MS017              1.   1MS023            .025
    1MUYD     1MS028             .25   1MS059            -.99
    1MUYD     1MS064              1.
    1MUYG     1MS002            -.04   1MS010            -.01
    1MUYG     1MS015              1.   1MS023            .025
    1MUYG     1MS028             .25   1MS035            -.99
    1MUYG     1MS064              1.
    SCCOL6    SCROW6             -1.   1RB045             -1.
    1P1BNP    1PD013             -1.   1PD025             50.
    1P1BNR    1PD018             -1.   1PD026             50.
    1P1SNP    1PD013              1.   1PD025            -50.
    1P1SNR    1PD018              1.   1PD026            -50.
    1P2BNP    1PD027             -1.   1PD031             -1.
    1P2BNP    1PD039             50.
    1P2BNR    1PD028             -1.   1PD032             -1.
    1P2BNR    1PD040             50.
    1P2SNP    1PD027              1.   1PD031              1.
    1P2SNP    1PD039            -50.
    1P2SNR    1PD028              1.   1PD032              1.
    1P2SNR    1PD040            -50.
    1PG11D    1DT061             -1.   1PD015             -1.
    1PG11I    1PD015             -1.
    1PG11R    1PD015             -1.   1RB040             -1.
    1PG11T    1PD015              1.   1PD019             -1.
    1PG11T    1PD023              1.
    1PG12D    1DT065             -1.   1PD016             -1.
    1PG12I    1PD016             -1.
    1PG12R    1PD016             -1.   1RB041             -1.
    1PG12T    1PD016              1.   1PD020             -1.
    1PG12T    1PD024              1.
    1PG1PD    1DT054             -1.   1PD013             -1.
    1PG1PD    1PD025             10.
    1PG1PI    1PD013             -1.   1PD025             10.
    1PG1PI    1PD069              1.
    1PG1PR    1PD013             -1.   1PD025             10.
    1PG1PR    1RB038             -1.
    1PG1PT    1PD013              1.   1PD017             -1.
    1PG1PT    1PD021              1.   1PD025            -10.
    1PG1RD    1DT058             -1.   1PD014             -1.
    1PG1RD    1PD026             10.
    1PG1RI    1PD014             -1.   1PD026             10.
    1PG1RI    1PD070              1.
    1PG1RR    1PD014             -1.   1PD026             10.
    1PG1RR    1RB039             -1.
    1PG1RT    1PD014              1.   1PD018             -1.
    1PG1RT    1PD022              1.   1PD026            -10.
    1PG21D    1DT061             -1.   1PD029             -1.
    1PG21I    1PD029             -1.
    1PG21M    1MS056             -1.   1PD029             -1.
    1PG21R    1PD029             -1.   1RB040             -1.
    1PG21T    1PD029              1.   1PD033             -1.
    1PG21T    1PD037              1.
    1PG22D    1DT065             -1.   1PD030             -1.
    1PG22I    1PD030             -1.
    1PG22M    1MS057             -1.   1PD030             -1.
    1PG22R    1PD030             -1.   1RB041             -1.
    1PG22T    1PD030              1.   1PD034             -1.
    1PG22T    1PD038              1.
    1PG2PD    1DT054             -1.   1PD027             -1.
    1PG2PD    1PD039             10.
    1PG2PI    1PD027             -1.   1PD039             10.
    1PG2PI    1PD071              1.
    1PG2PM    1MS042             -1.   1PD027             -1.
    1PG2PM    1PD039             10.
    1PG2PR    1PD027             -1.   1PD039             10.
    1PG2PR    1RB038             -1.
    1PG2PT    1PD027              1.   1PD031             -1.
    1PG2PT    1PD035              1.   1PD039            -10.
    1PG2RD    1DT058             -1.   1PD028             -1.
    1PG2RD    1PD040             10.
    1PG2RI    1PD028             -1.   1PD040             10.
    1PG2RI    1PD072              1.
    1PG2RM    1MS058             -1.   1PD028             -1.
    1PG2RM    1PD040             10.
    1PG2RR    1PD028             -1.   1PD040             10.
    1PG2RR    1RB039             -1.
    1PG2RT    1PD028              1.   1PD032             -1.
    1PG2RT    1PD036              1.   1PD040            -10.
    1PT11D    1DT061             -1.   1PD004           .1426
    1PT11D    1PD011              1.
    1PT11I    1PD011              1.   1PD019              1.
    1PT11R    1PD004           .2363   1PD011              1.
    1PT11R    1RB040             -1.
    1PT12D    1DT065             -1.   1PD004           .1426
    1PT12D    1PD012              1.
    1PT12I    1PD012              1.   1PD020              1.
    1PT12R    1PD004           .2363   1PD012              1.
    1PT12R    1RB041             -1.
    1PT1PD    1DT054             -1.   1PD004           .1426
    1PT1PD    1PD009              1.
    1PT1PI    1PD009              1.   1PD017              1.
    1PT1PR    1PD004           .2363   1PD009              1.
    1PT1PR    1RB038             -1.
    1PT1RD    1DT058             -1.   1PD004           .1426
    1PT1RD    1PD010              1.
    1PT1RI    1PD010              1.   1PD018              1.
    1PT1RR    1PD004           .2363   1PD010              1.
    1PT1RR    1RB039             -1.
    1PT21D    1DT061             -1.   1PD004           .2026
    1PT21D    1PD051              1.
    1PT21I    1PD019              1.   1PD051              1.
    1PT21R    1PD004           .2863   1PD051              1.
    1PT21R    1RB040             -1.
    1PT22D    1DT065             -1.   1PD004           .2026
    1PT22D    1PD052              1.
    1PT22I    1PD020              1.   1PD052              1.
    1PT22R    1PD004           .2863   1PD052              1.
    1PT22R    1RB041             -1.
    1PT2PD    1DT054             -1.   1PD004           .2026
    1PT2PD    1PD049              1.
    1PT2PI    1PD017              1.   1PD049              1.
    1PT2PR    1PD004           .2863   1PD049              1.
    1PT2PR    1RB038             -1.
    1PT2RD    1DT058             -1.   1PD004           .2026
    1PT2RD    1PD050              1.
    1PT2RI    1PD018              1.   1PD050              1.
    1PT2RR    1PD004           .2863   1PD050              1.
    1PT2RR    1RB039             -1.
    1PT31D    1DT061             -1.   1PD004           .1476
    1PT31D    1PD055              1.
    1PT31I    1PD019              1.   1PD055              1.
    1PT31R    1PD004           .3526   1PD055              1.
    1PT31R    1RB040             -1.
    1PT32D    1DT065             -1.   1PD004           .1476
    1PT32D    1PD056              1.
    1PT32I    1PD020              1.   1PD056              1.
    1PT32R    1PD004           .3526   1PD056              1.
    1PT32R    1RB041             -1.
    1PT3PD    1DT054             -1.   1PD004           .1476
    1PT3PD    1PD053              1.
    1PT3PI    1PD017              1.   1PD053              1.
    1PT3PR    1PD004           .3526   1PD053              1.
    1PT3PR    1RB038             -1.
    1PT3RD    1DT058             -1.   1PD004           .1476
    1PT3RD    1PD054              1.
    1PT3RI    1PD018              1.   1PD054              1.
    1PT3RR    1PD004           .3526   1PD054              1.
    1PT3RR    1RB039             -1.
    1PT41D    1DT061             -1.   1PD004           .0713
    1PT41D    1PD059              1.
    1PT41I    1PD019              1.   1PD059              1.
    1PT41R    1PD004           .2863   1PD059              1.
    1PT41R    1RB040             -1.
    1PT42D    1DT065             -1.   1PD004           .0713
    1PT42D    1PD060              1.
    1PT42I    1PD020              1.   1PD060              1.
    1PT42R    1PD004           .2863   1PD060              1.
    1PT42R    1RB041             -1.
    1PT4PD    1DT054             -1.   1PD004           .0713
    1PT4PD    1PD057              1.
    1PT4PI    1PD017              1.   1PD057              1.
    1PT4PR    1PD004           .2863   1PD057              1.
    1PT4PR    1RB038             -1.
    1PT4RD    1DT058             -1.   1PD004           .0713
    1PT4RD    1PD058              1.
    1PT4RI    1PD018              1.   1PD058              1.
    1PT4RR    1PD004           .2863   1PD058              1.
    1PT4RR    1RB039             -1.
    1PT51D    1DT061             -1.   1PD004            .395
    1PT51D    1PD063              1.
    1PT51I    1PD033              1.   1PD063              1.
    1PT51M    1MS056             -1.   1PD004           .1467
    1PT51M    1PD063              1.
    1PT51R    1PD004           .1239   1PD063              1.
    1PT51R    1RB040             -1.
    1PT52D    1DT065             -1.   1PD004           .4066
    1PT52D    1PD064              1.
    1PT52I    1PD034              1.   1PD064              1.
    1PT52M    1MS057             -1.   1PD004            .151
    1PT52M    1PD064              1.
    1PT52R    1PD004           .1273   1PD064              1.
    1PT52R    1RB041             -1.
    1PT5PD    1DT054             -1.   1PD004           .3536
    1PT5PD    1PD061              1.
    1PT5PI    1PD031              1.   1PD061              1.
    1PT5PM    1MS042             -1.   1PD004           .1313
    1PT5PM    1PD061              1.
    1PT5PR    1PD004            .112   1PD061              1.
    1PT5PR    1RB038             -1.
    1PT5RD    1DT058             -1.   1PD004           .3536
    1PT5RD    1PD062              1.
    1PT5RI    1PD032              1.   1PD062              1.
    1PT5RM    1MS058             -1.   1PD004           .1313
    1PT5RM    1PD062              1.
    1PT5RR    1PD004            .112   1PD062              1.
    1PT5RR    1RB039             -1.
    1PT61D    1DT061             -1.   1PD004           .3085
    1PT61D    1PD067              1.
    1PT61I    1PD033              1.   1PD067              1.
    1PT61M    1MS056             -1.   1PD004           .2547
    1PT61M    1PD067              1.
    1PT61R    1PD004           .3074   1PD067              1.
    1PT61R    1RB040             -1.
    1PT62D    1DT065             -1.   1PD004           .3175
    1PT62D    1PD068              1.
    1PT62I    1PD034              1.   1PD068              1.
    1PT62M    1MS057             -1.   1PD004           .2623
    1PT62M    1PD068              1.
    1PT62R    1PD004           .3161   1PD068              1.
    1PT62R    1RB041             -1.
    1PT6PD    1DT054             -1.   1PD004           .2766
    1PT6PD    1PD065              1.
    1PT6PI    1PD031              1.   1PD065              1.
    1PT6PM    1MS042             -1.   1PD004            .228
    1PT6PM    1PD065              1.
    1PT6PR    1PD004            .277   1PD065              1.
    1PT6PR    1RB038             -1.
    1PT6RD    1DT058             -1.   1PD004           .2766
    1PT6RD    1PD066              1.
    1PT6RI    1PD032              1.   1PD066              1.
    1PT6RM    1MS058             -1.   1PD004            .228
    1PT6RM    1PD066              1.
    1PT6RR    1PD004            .277   1PD066              1.
    1PT6RR    1RB039             -1.
    1PTCST    .COSTA              1.   1PD004             -1.
    1R1IN     1RB040             -1.   1RB091              1.
    1R2IN     1RB041             -1.   1RB092              1.
    1RALKI    1RB004            .588   1RB005             .65
    1RALKI    1RB031            .265   1RB037              1.
    1RALKI    1RB076          -.3208   1RB079              1.
    1RALKI    1RB088            6.01
    1RBGAS    1RB059           16.09   1RB060             -4.
    1RBGAS    1RB061            -86.   1RB062            -84.
    1RBGAS    1RB063           -80.4   1RB064           -84.5
    1RBGAS    1RB065           -82.5   1RB066           -79.4
    1RBGAS    1RB070              1.   1RB086              1.
    1RBGAS    1RB087           -3.57
    1RBIC4    1RB005             -1.   1RB087          -3.255
    1RBNC4    1RB006             -1.   1RB087          -3.045
    1RC3FU    1RB003              1.   1RB031          -3.836
    1RC4FU    1RB006              1.   1RB031          -4.325
    1RC6      1RB008              1.   1RB045              1.
    1RC6      1RB088              1.
    1RC7      1RB009              1.   1RB045              1.
    1RC7TX    1RB009              1.   1RB011             -1.
    1RC8TX    1RB010              1.   1RB011             -1.
    1RCOKE    1RB022              1.   1RB088            14.5
    1RCOKR    1RB001            -.41   1RB002          -.0072
    1RCOKR    1RB003          -.0183   1RB004          -.0106
    1RCOKR    1RB005          -.0032   1RB006          -.0116
    1RCOKR    1RB017              1.   1RB018           -.125
    1RCOKR    1RB019             -.3   1RB020            -.34
    1RCOKR    1RB022          -.0427   1RB031            .221
    1RCOKR    1RB036              1.   1RB076          -.0582
    1RCOKS    .COSTA             10.   1RB001           -.492
    1RCOKS    1RB002           -.056   1RB003           -.033
    1RCOKS    1RB004           -.075   1RB005           -.037
    1RCOKS    1RB006           -.019   1RB018           -.015
    1RCOKS    1RB019           -.015   1RB023          -.0399
    1RCOKS    1RB024           -.024   1RB025           -.011
    1RCOKS    1RB026           -.035   1RB027            .538
    1RCOKS    1RB031            .273   1RB036              1.
    1RCOKS    1RB076           -.181
    1RCOKT    .COSTA             10.   1RB001           -.572
    1RCOKT    1RB021           -.423   1RB023          -.0651
    1RCOKT    1RB030              1.   1RB031            .228
    1RCOKT    1RB036              1.   1RB076           -.065
    1RCRUD    1RB001          -.0005   1RB003          -.0083
    1RCRUD    1RB005           -.001   1RB006           -.001
    1RCRUD    1RB007            -.07   1RB008          -.0375
    1RCRUD    1RB009          -.0375   1RB010           -.037
    1RCRUD    1RB011            -.14   1RB012          -.0451
    1RCRUD    1RB013            -.11   1RB014            -.07
    1RCRUD    1RB015           -.023   1RB016           -.325
    1RCRUD    1RB017           -.139   1RB031            .124
    1RCRUD    1RB033              1.   1RB076          -.0178
    1RCRUD    1RB087           -3.01
    1RDLCO    1RB026              1.   1RB088              1.
    1REXTR    1RB083              1.   1RB084              1.
    1REXTR    1RB088             6.3
    1RF2GO    1RB015              1.   1RB041              1.
    1RF2GO    1RB067           -14.6   1RB073              1.
    1RF2KR    1RB013              1.   1RB041              1.
    1RF2KR    1RB067            -1.6   1RB073             -.5
    1RF3CO    1RB026              1.   1RB042              1.
    1RF6CO    1RB026              1.   1RB043              1.
    1RF6CO    1RB068           -44.5   1RB077            43.7
    1RF6CO    1RB088            2.52
    1RF6SL    1RB027              1.   1RB043              1.
    1RF6SL    1RB068           -12.7   1RB077            11.9
    1RF6SL    1RB088            2.52
    1RFC01    1RB001           -.374   1RB002           -.065
    1RFC01    1RB003          -.0382   1RB004          -.0753
    1RFC01    1RB005          -.0535   1RB006           -.018
    1RFC01    1RB016              1.   1RB024           -.422
    1RFC01    1RB025           -.181   1RB026           -.144
    1RFC01    1RB027            -.05   1RB031            .159
    1RFC01    1RB032          .06061   1RB076          -.0783
    1RFC01    1RB080              1.
    1RFC02    1RB001           -.333   1RB002          -.0595
    1RFC02    1RB003          -.0293   1RB004          -.0724
    1RFC02    1RB005          -.0455   1RB006          -.0153
    1RFC02    1RB016              1.   1RB024           -.411
    1RFC02    1RB025           -.176   1RB026           -.166
    1RFC02    1RB027            -.06   1RB031            .152
    1RFC02    1RB032          .04662   1RB076          -.0739
    1RFC02    1RB080              1.
    1RFC03    1RB001           -.335   1RB002          -.0603
    1RFC03    1RB003          -.0313   1RB004          -.0726
    1RFC03    1RB005           -.051   1RB006          -.0167
    1RFC03    1RB015            .073   1RB016            .927
    1RFC03    1RB024           -.419   1RB025            -.18
    1RFC03    1RB026           -.144   1RB027           -.061
    1RFC03    1RB031            .154   1RB032          .05618
    1RFC03    1RB076          -.0763   1RB080              1.
    1RFC04    1RB001           -.363   1RB002           -.061
    1RFC04    1RB003           -.031   1RB004           -.073
    1RFC04    1RB005          -.0457   1RB006          -.0154
    1RFC04    1RB015            .057   1RB016            .943
    1RFC04    1RB024           -.398   1RB025            -.17
    1RFC04    1RB026           -.179   1RB027            -.06
    1RFC04    1RB031            .153   1RB032          .04396
    1RFC04    1RB076          -.0763   1RB080              1.
    1RFC05    1RB001            -.32   1RB002          -.0592
    1RFC05    1RB003          -.0275   1RB004          -.0685
    1RFC05    1RB005          -.0407   1RB006          -.0139
    1RFC05    1RB016            .834   1RB020            .166
    1RFC05    1RB024           -.402   1RB025           -.173
    1RFC05    1RB026           -.175   1RB027            -.06
    1RFC05    1RB031            .149   1RB032          .05056
    1RFC05    1RB076          -.0767   1RB080              1.
    1RFC06    1RB001           -.402   1RB002          -.0563
    1RFC06    1RB003           -.031   1RB004          -.0718
    1RFC06    1RB005            -.04   1RB006          -.0142
    1RFC06    1RB016            .867   1RB020            .133
    1RFC06    1RB024           -.384   1RB025           -.165
    1RFC06    1RB026           -.204   1RB027           -.058
    1RFC06    1RB031             .15   1RB032          .04044
    1RFC06    1RB076           -.074   1RB080              1.
    1RFC07    1RB001           -.388   1RB002          -.0622
    1RFC07    1RB003          -.0306   1RB004          -.0707
    1RFC07    1RB005          -.0426   1RB006          -.0145
    1RFC07    1RB015            .062   1RB016            .783
    1RFC07    1RB020            .155   1RB024           -.402
    1RFC07    1RB025           -.173   1RB026           -.158
    1RFC07    1RB027           -.064   1RB031            .152
    1RFC07    1RB032          .04744   1RB076          -.0743
    1RFC07    1RB080              1.
    1RFC08    1RB001           -.431   1RB002           -.058
    1RFC08    1RB003           -.032   1RB004          -.0734
    1RFC08    1RB005          -.0394   1RB006          -.0142
    1RFC08    1RB015             .05   1RB016            .824
    1RFC08    1RB020            .126   1RB024           -.372
    1RFC08    1RB025           -.159   1RB026           -.197
    1RFC08    1RB027           -.075   1RB031            .151
    1RFC08    1RB032          .03842   1RB076          -.0741
    1RFC08    1RB080              1.
    1RFC09    1RB001            -.31   1RB002          -.0615
    1RFC09    1RB003          -.0293   1RB004          -.0689
    1RFC09    1RB005          -.0388   1RB006          -.0136
    1RFC09    1RB016            .928   1RB021            .078
    1RFC09    1RB024           -.415   1RB025           -.178
    1RFC09    1RB026           -.169   1RB027            -.06
    1RFC09    1RB031             .15   1RB032          .05624
    1RFC09    1RB076          -.0748   1RB080              1.
    1RFC10    1RB001            -.33   1RB002          -.0601
    1RFC10    1RB003          -.0288   1RB004          -.0693
    1RFC10    1RB005          -.0384   1RB006          -.0134
    1RFC10    1RB016            .944   1RB021            .056
    1RFC10    1RB024           -.397   1RB025            -.17
    1RFC10    1RB026           -.193   1RB027           -.065
    1RFC10    1RB031            .149   1RB032          .04399
    1RFC10    1RB076          -.0724   1RB080              1.
    1RFC11    1RB001            -.29   1RB002          -.0585
    1RFC11    1RB003          -.0262   1RB004          -.0668
    1RFC11    1RB005          -.0384   1RB006          -.0134
    1RFC11    1RB015            .068   1RB016            .865
    1RFC11    1RB021            .067   1RB024           -.406
    1RFC11    1RB025           -.174   1RB026            -.18
    1RFC11    1RB027           -.065   1RB031            .148
    1RFC11    1RB032          .05241   1RB076          -.0733
    1RFC11    1RB080              1.
    1RFC12    1RB001           -.385   1RB002          -.0582
    1RFC12    1RB003          -.0304   1RB004          -.0738
    1RFC12    1RB005          -.0401   1RB006          -.0145
    1RFC12    1RB015            .054   1RB016            .893
    1RFC12    1RB021            .053   1RB024           -.387
    1RFC12    1RB025           -.166   1RB026           -.201
    1RFC12    1RB027            -.06   1RB031            .151
    1RFC12    1RB032          .04161   1RB076          -.0734
    1RFC12    1RB080              1.
    1RFC13    1RB001           -.263   1RB002          -.0519
    1RFC13    1RB003          -.0228   1RB004          -.0585
    1RFC13    1RB005          -.0301   1RB006          -.0108
    1RFC13    1RB016            .783   1RB020            .156
    1RFC13    1RB021            .061   1RB024           -.387
    1RFC13    1RB025           -.166   1RB026           -.217
    1RFC13    1RB027           -.083   1RB031            .141
    1RFC13    1RB032          .04748   1RB076          -.0699
    1RFC13    1RB080              1.
    1RFC14    1RB001             -.4   1RB002          -.0558
    1RFC14    1RB003          -.0292   1RB004          -.0694
    1RFC14    1RB005          -.0332   1RB006          -.0124
    1RFC14    1RB016            .825   1RB020            .126
    1RFC14    1RB021            .049   1RB024           -.372
    1RFC14    1RB025           -.159   1RB026           -.211
    1RFC14    1RB027           -.084   1RB031            .147
    1RFC14    1RB032          .03845   1RB076          -.0713
    1RFC14    1RB080              1.
    1RFC15    1RB001           -.356   1RB002          -.0593
    1RFC15    1RB003          -.0276   1RB004          -.0652
    1RFC15    1RB005          -.0331   1RB006          -.0118
    1RFC15    1RB015            .058   1RB016            .738
    1RFC15    1RB020            .147   1RB021            .057
    1RFC15    1RB024           -.385   1RB025           -.165
    1RFC15    1RB026           -.205   1RB027            -.07
    1RFC15    1RB031            .146   1RB032          .04472
    1RFC15    1RB076          -.0728   1RB080              1.
    1RFC16    1RB001           -.436   1RB002          -.0578
    1RFC16    1RB003          -.0302   1RB004          -.0724
    1RFC16    1RB005          -.0339   1RB006          -.0126
    1RFC16    1RB015            .048   1RB016            .785
    1RFC16    1RB020             .12   1RB021            .047
    1RFC16    1RB024           -.364   1RB025           -.156
    1RFC16    1RB026           -.186   1RB027           -.107
    1RFC16    1RB031            .148   1RB032          .03662
    1RFC16    1RB076          -.0719   1RB080              1.
    1RFC17    1RB001           -.404   1RB002          -.0625
    1RFC17    1RB003          -.0318   1RB004          -.0716
    1RFC17    1RB005          -.0409   1RB006          -.0141
    1RFC17    1RB015            .027   1RB016            .857
    1RFC17    1RB020            .116   1RB024           -.389
    1RFC17    1RB025           -.167   1RB026           -.192
    1RFC17    1RB027           -.058   1RB031            .152
    1RFC17    1RB032          .04202   1RB076          -.0747
    1RFC17    1RB080              1.
    1RFC18    1RB001           -.385   1RB002          -.0615
    1RFC18    1RB003          -.0307   1RB004          -.0713
    1RFC18    1RB005          -.0405   1RB006           -.014
    1RFC18    1RB016            .881   1RB020            .119
    1RFC18    1RB024           -.396   1RB025            -.17
    1RFC18    1RB026           -.189   1RB027           -.056
    1RFC18    1RB031            .152   1RB032           .0432
    1RFC18    1RB076          -.0746   1RB080              1.
    1RFC19    1RB001             -.4   1RB002          -.0617
    1RFC19    1RB003          -.0316   1RB004          -.0708
    1RFC19    1RB005          -.0399   1RB006          -.0137
    1RFC19    1RB015            .027   1RB016            .843
    1RFC19    1RB020             .13   1RB024           -.388
    1RFC19    1RB025           -.166   1RB026           -.198
    1RFC19    1RB027           -.058   1RB031            .151
    1RFC19    1RB032          .04132   1RB076          -.0744
    1RFC19    1RB080              1.
    1RFC20    1RB001           -.382   1RB002          -.0606
    1RFC20    1RB003          -.0307   1RB004          -.0701
    1RFC20    1RB005          -.0398   1RB006          -.0138
    1RFC20    1RB016            .866   1RB020            .134
    1RFC20    1RB024           -.393   1RB025           -.169
    1RFC20    1RB026           -.195   1RB027           -.056
    1RFC20    1RB031             .15   1RB032          .04246
    1RFC20    1RB076          -.0746   1RB080              1.
    1RFC21    1RB001           -.368   1RB002          -.0541
    1RFC21    1RB003          -.0282   1RB004          -.0676
    1RFC21    1RB005          -.0339   1RB006          -.0123
    1RFC21    1RB015            .025   1RB016            .801
    1RFC21    1RB020            .124   1RB021             .05
    1RFC21    1RB024            -.37   1RB025           -.159
    1RFC21    1RB026           -.214   1RB027           -.086
    1RFC21    1RB031            .146   1RB032          .03925
    1RFC21    1RB076          -.0717   1RB080              1.
    1RFC22    1RB001            -.36   1RB002          -.0539
    1RFC22    1RB003           -.028   1RB004          -.0672
    1RFC22    1RB005          -.0343   1RB006          -.0125
    1RFC22    1RB016            .821   1RB020            .127
    1RFC22    1RB021            .052   1RB024           -.375
    1RFC22    1RB025           -.161   1RB026           -.216
    1RFC22    1RB027           -.079   1RB031            .146
    1RFC22    1RB032          .04027   1RB076          -.0719
    1RFC22    1RB080              1.
    1RFCCU    1RB080             -1.
    1RFCOD    1RB019              1.   1RB031            .053
    1RFCOD    1RB035              1.   1RB041              1.
    1RFCOD    1RB067             3.4   1RB076           -.008
    1RFDIE    1RB014              1.   1RB041              1.
    1RFDIE    1RB067            -8.6
    1RFHDO    1RB019            .385   1RB026            .615
    1RFHDO    1RB035              1.   1RB042              1.
    1RFHDO    1RB076           -.008
    1RFKER    1RB013              1.   1RB040              1.
    1RFLCO    1RB026              1.   1RB031            .053
    1RFLCO    1RB035              1.   1RB041              1.
    1RFLCO    1RB067            15.4   1RB076           -.008
    1RFRED    1RB041              1.   1RB067             5.4
    1RFRED    1RB085              1.
    1RFRES    1RB017              1.   1RB043              1.
    1RFRES    1RB068            10.9   1RB077           -11.7
    1RFRES    1RB088            2.52
    1RFRHC    1RB028              1.   1RB042              1.
    1RFRHC    1RB059            -.01   1RB060              4.
    1RFRHC    1RB061            88.5   1RB062            87.5
    1RFRHC    1RB063            85.7   1RB064            76.3
    1RFRHC    1RB065            75.6   1RB066            74.5
    1RFRHC    1RB070             -1.
    1RFTHR    1RB030              1.   1RB043              1.
    1RFTHR    1RB068             -.6   1RB077             -.2
    1RFTHR    1RB088            2.52
    1RGAS     1RB001              1.   1RB031             -1.
    1RLFUL    1RB026              1.   1RB031            -6.5
    1RLFUL    1RB078           -44.5
    1RLPG     1RB003              1.   1RB044              1.
    1RLPG     1RB088            1.89
    1RMCST    .COSTA              1.   1RB087              1.
    1ROCST    .COSTA              1.   1RB076              1.
    1RP84R    1RB001           -.629   1RB003          -.0219
    1RP84R    1RB005          -.0045   1RB006           -.001
    1RP84R    1RB011           1.119   1RB031             .37
    1RP84R    1RB034           1.119   1RB048            3.84
    1RP84R    1RB049             8.9   1RB050             -4.
    1RP84R    1RB051           -92.5   1RB052           -91.4
    1RP84R    1RB053           -88.9   1RB054            -84.
    1RP84R    1RB055           -83.1   1RB056           -80.4
    1RP84R    1RB069              1.   1RB076          -.0379
    1RP84R    1RB081           -92.9   1RB082           -85.1
    1RP88R    1RB001           -.756   1RB003          -.0284
    1RP88R    1RB005          -.0061   1RB006          -.0013
    1RP88R    1RB011           1.146   1RB031            .405
    1RP88R    1RB034           1.146   1RB048            4.83
    1RP88R    1RB049             7.5   1RB050             -4.
    1RP88R    1RB051           -94.5   1RB052           -93.7
    1RP88R    1RB053           -91.7   1RB054           -85.5
    1RP88R    1RB055           -84.7   1RB056           -82.5
    1RP88R    1RB069              1.   1RB076          -.0458
    1RP88R    1RB081           -94.7   1RB082           -86.5
    1RP92R    1RB001           -.898   1RB003          -.0359
    1RP92R    1RB005           -.008   1RB006          -.0016
    1RP92R    1RB011           1.182   1RB031            .455
    1RP92R    1RB034           1.182   1RB048            6.12
    1RP92R    1RB049             6.2   1RB050             -4.
    1RP92R    1RB051           -96.5   1RB052           -95.9
    1RP92R    1RB053           -94.4   1RB054           -86.9
    1RP92R    1RB055           -86.4   1RB056           -84.5
    1RP92R    1RB069              1.   1RB076          -.0592
    1RP92R    1RB081           -96.6   1RB082           -87.7
    1RP95R    1RB001          -1.056   1RB003          -.0433
    1RP95R    1RB005            -.01   1RB006          -.0019
    1RP95R    1RB011           1.224   1RB031             .51
    1RP95R    1RB034           1.224   1RB048            7.62
    1RP95R    1RB049             5.2   1RB050             -4.
    1RP95R    1RB051            -98.   1RB052           -97.5
    1RP95R    1RB053           -96.3   1RB054            -88.
    1RP95R    1RB055           -87.7   1RB056           -86.1
    1RP95R    1RB069              1.   1RB076           -.086
    1RP95R    1RB081           -98.1   1RB082           -88.9
    1RPALC    1RB083              1.   1RB087             -5.
    1RPALK    1RB004            .588   1RB005             .65
    1RPALK    1RB031            .265   1RB037              1.
    1RPALK    1RB048            4.39   1RB049            -7.3
    1RPALK    1RB050             -4.   1RB051          -102.3
    1RPALK    1RB052          -101.1   1RB053           -98.8
    1RPALK    1RB054            -96.   1RB055           -94.9
    1RPALK    1RB056           -91.5   1RB069              1.
    1RPALK    1RB076          -.3208   1RB081          -104.6
    1RPALK    1RB082          -102.3
    1RPBBL    1RB038              1.   1RB046            95.5
    1RPBBL    1RB047             99.   1RB048            -10.
    1RPBBL    1RB049             20.   1RB069             -1.
    1RPBBL    1RB071             -.1   1RB072             .25
    1RPBBL    1RB076          -.0139
    1RPCAT    1RB024              1.   1RB048            9.36
    1RPCAT    1RB049           -32.2   1RB050             -4.
    1RPCAT    1RB051            -98.   1RB052           -97.5
    1RPCAT    1RB053           -96.2   1RB054           -80.7
    1RPCAT    1RB055           -80.5   1RB056           -79.2
    1RPCAT    1RB069              1.   1RB072             -1.
    1RPCAT    1RB076           -.025   1RB081           -98.1
    1RPCAT    1RB082           -81.3
    1RPIC4    1RB005              1.   1RB048           95.62
    1RPIC4    1RB049           -130.   1RB050             -4.
    1RPIC4    1RB051           -107.   1RB052          -105.8
    1RPIC4    1RB053          -103.5   1RB054          -106.7
    1RPIC4    1RB055          -105.6   1RB056          -103.1
    1RPIC4    1RB069              1.   1RB081          -106.9
    1RPIC4    1RB082          -107.3
    1RPIN     1RB038             -1.   1RB089              1.
    1RPMON    1RB046             -.5   1RB054              1.
    1RPMON    1RB055              1.   1RB056              1.
    1RPMON    1RB082              1.
    1RPNAG    1RB086              1.   1RB087             -5.
    1RPNC4    1RB006              1.   1RB048           68.56
    1RPNC4    1RB049           -130.   1RB050             -4.
    1RPNC4    1RB051            -97.   1RB052           -96.4
    1RPNC4    1RB053           -94.8   1RB054           -96.7
    1RPNC4    1RB055           -96.2   1RB056           -94.4
    1RPNC4    1RB069              1.   1RB081           -97.2
    1RPNC4    1RB082           -97.6
    1RPREV    .COSTA             -1.   1RB088             -1.
    1RPRON    1RB046             -.5   1RB047             -1.
    1RPRON    1RB051              1.   1RB052              1.
    1RPRON    1RB053              1.   1RB081              1.
    1RPTEL    1RB050              1.   1RB051             -1.
    1RPTEL    1RB052            -1.3   1RB053            -2.5
    1RPTEL    1RB054            -1.1   1RB055            -1.4
    1RPTEL    1RB056            -2.9   1RB081             -.9
    1RPTEL    1RB082             -.8   1RB087          -.0886
    1RPUFU    1RB031             -1.   1RB087            -.35
    1RR80R    1RB001           -.517   1RB003          -.0164
    1RR80R    1RB005          -.0032   1RB006          -.0007
    1RR80R    1RB011           1.094   1RB031             .35
    1RR80R    1RB034           1.094   1RB059            2.89
    1RR80R    1RB060             -4.   1RB061           -90.2
    1RR80R    1RB062           -88.5   1RB063           -85.4
    1RR80R    1RB064            -83.   1RB065           -81.3
    1RR80R    1RB066           -78.5   1RB070              1.
    1RR80R    1RB076          -.0327
    1RR84R    1RB001           -.629   1RB003          -.0219
    1RR84R    1RB005          -.0045   1RB006           -.001
    1RR84R    1RB011           1.119   1RB031             .37
    1RR84R    1RB034           1.119   1RB059            3.72
    1RR84R    1RB060             -4.   1RB061           -92.2
    1RR84R    1RB062           -90.8   1RB063           -88.1
    1RR84R    1RB064           -84.6   1RB065           -83.3
    1RR84R    1RB066           -80.8   1RB070              1.
    1RR84R    1RB076          -.0379
    1RR88R    1RB001           -.756   1RB003          -.0284
    1RR88R    1RB005          -.0061   1RB006          -.0013
    1RR88R    1RB011           1.146   1RB031            .405
    1RR88R    1RB034           1.146   1RB059            4.69
    1RR88R    1RB060             -4.   1RB061           -94.2
    1RR88R    1RB062           -93.1   1RB063           -90.9
    1RR88R    1RB064           -86.1   1RB065           -84.9
    1RR88R    1RB066           -82.9   1RB070              1.
    1RR88R    1RB076          -.0458
    1RR92R    1RB001           -.898   1RB003          -.0359
    1RR92R    1RB005           -.008   1RB006          -.0016
    1RR92R    1RB011           1.182   1RB031            .455
    1RR92R    1RB034           1.182   1RB059            5.93
    1RR92R    1RB060             -4.   1RB061           -96.2
    1RR92R    1RB062           -95.3   1RB063           -93.6
    1RR92R    1RB064           -87.5   1RB065           -86.6
    1RR92R    1RB066           -84.9   1RB070              1.
    1RR92R    1RB076          -.0592
    1RR95R    1RB001          -1.056   1RB003          -.0433
    1RR95R    1RB005            -.01   1RB006          -.0019
    1RR95R    1RB011           1.224   1RB031             .51
    1RR95R    1RB034           1.224   1RB059            7.39
    1RR95R    1RB060             -4.   1RB061           -97.7
    1RR95R    1RB062           -96.9   1RB063           -95.5
    1RR95R    1RB064           -88.6   1RB065           -87.9
    1RR95R    1RB066           -86.5   1RB070              1.
    1RR95R    1RB076           -.086
    1RRALK    1RB004            .588   1RB005             .65
    1RRALK    1RB031            .265   1RB037              1.
    1RRALK    1RB059            4.26   1RB060             -4.
    1RRALK    1RB061          -103.7   1RB062          -102.1
    1RRALK    1RB063           -99.4   1RB064          -101.3
    1RRALK    1RB065           -99.4   1RB066           -96.6
    1RRALK    1RB070              1.   1RB076          -.3208
    1RRBBL    1RB039              1.   1RB057            91.2
    1RRBBL    1RB058             93.   1RB059            -10.
    1RRBBL    1RB070             -1.   1RB076          -.0046
    1RRC3P    1RB002          1.5873   1RB003         -.15873
    1RRC3P    1RB031            .075   1RB059            4.03
    1RRC3P    1RB060             -4.   1RB061          -101.7
    1RRC3P    1RB062           -101.   1RB063           -99.7
    1RRC3P    1RB064           -79.3   1RB065           -78.6
    1RRC3P    1RB066           -77.5   1RB070              1.
    1RRC3P    1RB076          -.2695
    1RRC45    1RB007              1.   1RB059           26.57
    1RRC45    1RB060             -4.   1RB061            -88.
    1RRC45    1RB062            -86.   1RB063           -82.2
    1RRC45    1RB064            -89.   1RB065           -86.8
    1RRC45    1RB066           -82.9   1RB070              1.
    1RRC4O    1RB004              1.   1RB059           69.09
    1RRC4O    1RB060             -4.   1RB061          -106.7
    1RRC4O    1RB062          -105.2   1RB063          -102.7
    1RRC4O    1RB064           -92.3   1RB065           -90.8
    1RRC4O    1RB066           -88.4   1RB070              1.
    1RRC4P    1RB004            1.25   1RB031            .075
    1RRC4P    1RB059            4.03   1RB060             -4.
    1RRC4P    1RB061          -101.7   1RB062           -101.
    1RRC4P    1RB063           -99.7   1RB064           -79.3
    1RRC4P    1RB065           -78.6   1RB066           -77.5
    1RRC4P    1RB070              1.   1RB076          -.2122
    1RRC6     1RB008              1.   1RB059            4.21
    1RRC6     1RB060             -4.   1RB061           -81.2
    1RRC6     1RB062            -79.   1RB063           -75.1
    1RRC6     1RB064           -80.3   1RB065           -77.9
    1RRC6     1RB066            -74.   1RB070              1.
    1RRC7     1RB009              1.   1RB059             2.1
    1RRC7     1RB060             -4.   1RB061           -76.2
    1RRC7     1RB062           -74.2   1RB063           -70.7
    1RRC7     1RB064           -76.1   1RB065           -74.1
    1RRC7     1RB066            -71.   1RB070              1.
    1RRC8E    1RB010              1.   1RB059             .97
    1RRC8E    1RB060             -4.   1RB061           -74.5
    1RRC8E    1RB062           -72.5   1RB063           -67.4
    1RRC8E    1RB064           -74.6   1RB065           -72.4
    1RRC8E    1RB066           -67.4   1RB070              1.
    1RRCAT    1RB025              1.   1RB028             -.5
    1RRCAT    1RB059             .08   1RB060             -4.
    1RRCAT    1RB061           -93.7   1RB062            -93.
    1RRCAT    1RB063           -91.4   1RB064           -81.3
    1RRCAT    1RB065           -80.7   1RB066           -79.6
    1RRCAT    1RB070              1.
    1RRCOK    1RB018              1.   1RB059            6.62
    1RRCOK    1RB060             -4.   1RB061           -84.7
    1RRCOK    1RB062           -82.6   1RB063           -79.1
    1RRCOK    1RB064           -74.6   1RB065           -73.7
    1RRCOK    1RB066           -72.3   1RB070              1.
    1RRCTF    1RB028              1.   1RB041              1.
    1RRCTF    1RB059            -.01   1RB060              4.
    1RRCTF    1RB061            88.5   1RB062            87.5
    1RRCTF    1RB063            85.7   1RB064            76.3
    1RRCTF    1RB065            75.6   1RB066            74.5
    1RRCTF    1RB067             23.   1RB070             -1.
    1RRDEX    1RB026              1.   1RB076            -.22
    1RRDEX    1RB084            -.26   1RB085            -.74
    1RRHN     1RB011              1.   1RB059             .38
    1RRHN     1RB060             -4.   1RB061           -66.7
    1RRHN     1RB062           -62.8   1RB063           -56.3
    1RRHN     1RB064           -67.3   1RB065           -63.3
    1RRHN     1RB066           -56.8   1RB070              1.
    1RRIC4    1RB005              1.   1RB059           92.73
    1RRIC4    1RB060             -4.   1RB061          -106.7
    1RRIC4    1RB062          -105.2   1RB063          -102.7
    1RRIC4    1RB064          -107.3   1RB065          -105.8
    1RRIC4    1RB066          -103.5   1RB070              1.
    1RRIN     1RB039             -1.   1RB090              1.
    1RRLCT    1RB024              1.   1RB059            9.08
    1RRLCT    1RB060             -4.   1RB061           -97.7
    1RRLCT    1RB062           -96.9   1RB063           -95.4
    1RRLCT    1RB064           -81.3   1RB065           -80.7
    1RRLCT    1RB066           -79.6   1RB070              1.
    1RRMON    1RB057           -.333   1RB064              1.
    1RRMON    1RB065              1.   1RB066              1.
    1RRNC4    1RB006              1.   1RB059           66.48
    1RRNC4    1RB060             -4.   1RB061           -96.7
    1RRNC4    1RB062           -95.8   1RB063            -94.
    1RRNC4    1RB064           -97.3   1RB065           -96.4
    1RRNC4    1RB066           -94.8   1RB070              1.
    1RRRON    1RB057           -.667   1RB058             -1.
    1RRRON    1RB061              1.   1RB062              1.
    1RRRON    1RB063              1.
    1RRTEL    1RB060              1.   1RB061            -1.1
    1RRTEL    1RB062            -1.6   1RB063            -3.3
    1RRTEL    1RB064             -.9   1RB065            -1.3
    1RRTEL    1RB066            -2.5   1RB087           .0886
    1RRTHR    1RB029              1.   1RB059            6.62
    1RRTHR    1RB060             -4.   1RB061           -78.7
    1RRTHR    1RB062           -76.2   1RB063           -72.1
    1RRTHR    1RB064           -74.6   1RB065            -72.
    1RRTHR    1RB066           -67.9   1RB070              1.
    1RRXF1    1RB011             .85   1RB012              1.
    1RRXF1    1RB013             -1.   1RB014             .15
    1RSBT     1RB045              1.   1RB087             -5.
    1RSCOK    1RB023              1.   1RB088             40.
    1RSLFU    1RB027              1.   1RB031            -6.5
    1RSLFU    1RB078           -12.7
    1RSNO1    1RB040              1.   1RB087            -4.2
    1RSNO2    1RB041              1.   1RB087           -3.78
    1RSNO3    1RB042              1.   1RB087           -3.57
    1RSPRM    1RB038              1.   1RB087           -5.46
    1RSREG    1RB039              1.   1RB087           -4.62
    1RTCHG    1RB001           -.359   1RB002           -.009
    1RTCHG    1RB003          -.0229   1RB004          -.0132
    1RTCHG    1RB005          -.0039   1RB006          -.0145
    1RTCHG    1RB027              1.   1RB029          -.2896
    1RTCHG    1RB030           -.639   1RB031              .4
    1RTCHG    1RB076           -.099
    1RTFUL    1RB030              1.   1RB031            -6.5
    1RTFUL    1RB078             -.6
    1RVFUL    1RB017              1.   1RB031            -6.5
    1RVFUL    1RB078            10.9
    1RXSFL    1RB031              1.   1RB088             .25
RHS
    ZZZZ0001  1DT058            480.   1DT061             30.
    ZZZZ0001  1DT065            210.   1DT054            120.
    ZZZZ0001  1MS042            15.3   1MS056              3.
    ZZZZ0001  1MS057             30.   1MS058             65.
    ZZZZ0001  1RB039           1050.   1RB040            150.
    ZZZZ0001  1RB041            540.   1RB038            243.
    ZZZZ0001  1DT017             .75   1DT027            37.5
    ZZZZ0001  1DT028           1200.   1DT029            480.
    ZZZZ0001  1DT030            480.   1DT031            105.
    ZZZZ0001  1DT032            135.   1DT033            450.
    ZZZZ0001  1DT038            225.   1DT039            300.
    ZZZZ0001  1DT040            480.   1DT041             45.
    ZZZZ0001  1DT045            570.   1DT047           112.5
    ZZZZ0001  1DT048             45.   1DT049            67.5
    ZZZZ0001  1DT050              3.   1DT068             45.
    ZZZZ0001  1DT072             30.   1DT074            120.
    ZZZZ0001  1DT075           13.34   1DT076            180.
    ZZZZ0001  1DT077             45.   1DT095            100.
    ZZZZ0001  1DT096            94.5   1DT097           465.5
    ZZZZ0001  1DT098           249.1   1DT099           648.2
    ZZZZ0001  1MS037             15.   1MS038           34.16
    ZZZZ0001  1MS039             4.5   1MS041           21.93
    ZZZZ0001  1MS060            147.   1MS061            270.
    ZZZZ0001  1MS062             72.   1MS063            180.
    ZZZZ0001  1MS064             60.   1MS075             45.
    ZZZZ0001  1MS076            118.   1MS077           125.5
    ZZZZ0001  1MS078            105.   1PD072             50.
    ZZZZ0001  1PD009             50.   1PD010            150.
    ZZZZ0001  1PD011             25.   1PD012             75.
    ZZZZ0001  1PD015             25.   1PD016            100.
    ZZZZ0001  1PD017            -25.   1PD018           -100.
    ZZZZ0001  1PD019            -25.   1PD020           -100.
    ZZZZ0001  1PD021            84.9   1PD022           203.5
    ZZZZ0001  1PD023            89.8   1PD024           219.2
    ZZZZ0001  1PD029             25.   1PD030             30.
    ZZZZ0001  1PD031            -25.   1PD032            -50.
    ZZZZ0001  1PD033            -25.   1PD034            -30.
    ZZZZ0001  1PD035             44.   1PD036            91.7
    ZZZZ0001  1PD037            69.5   1PD038             55.
    ZZZZ0001  1PD049             50.   1PD050            200.
    ZZZZ0001  1PD051             25.   1PD052             75.
    ZZZZ0001  1PD053             25.   1PD054            100.
    ZZZZ0001  1PD055             20.   1PD056             50.
    ZZZZ0001  1PD057             25.   1PD058            100.
    ZZZZ0001  1PD059             10.   1PD060             40.
    ZZZZ0001  1PD061             50.   1PD062            200.
    ZZZZ0001  1PD063             50.   1PD064            150.
    ZZZZ0001  1PD065             10.   1PD066             50.
    ZZZZ0001  1PD067             20.   1PD068             50.
    ZZZZ0001  1PD069             25.   1PD070            100.
    ZZZZ0001  1PD071             25.   1RB032              1.
    ZZZZ0001  1RB033           1800.   1RB034            315.
    ZZZZ0001  1RB035            195.   1RB036            360.
    ZZZZ0001  1RB037            105.   1RB042             30.
    ZZZZ0001  1RB043             90.   1RB044             90.
    ZZZZ0001  1RB083             15.   1RB086             75.
    ZZZZ0001  1RB089           721.5   1RB090          1626.7
    ZZZZ0001  1RB091            692.   1RB092          1229.7
ENDATA
