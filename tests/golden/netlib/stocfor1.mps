NAME          STOCFOR1 (STOCHFOR)
ROWS
 N  HARV
 L  BOUND301
 L  BOUND401
 L  BOUND501
 L  BOUND601
 L  BOUND701
 L  BOUND801
 E  REGEN101
 E  REGEN201
 E  REGEN301
 E  REGEN401
 E  REGEN501
 E  REGEN601
 E  REGEN701
 E  REGEN801
 E  YIELD101
 L  BOUND302
 L  BOUND402
 L  BOUND502
 L  BOUND602
 L  BOUND702
 L  BOUND802
 E  REGEN102
 E  REGEN202
 E  REGEN302
 E  REGEN402
 E  REGEN502
 E  REGEN602
 E  REGEN702
 E  REGEN802
 L  TFLOW102
 G  TFLOW202
 E  YIELD202
 L  BOUND303
 L  BOUND403
 L  BOUND503
 L  BOUND603
 L  BOUND703
 L  BOUND803
 E  REGEN103
 E  REGEN203
 E  REGEN303
 E  REGEN403
 E  REGEN503
 E  REGEN603
 E  REGEN703
 E  REGEN803
 L  TFLOW103
 G  TFLOW203
 E  YIELD303
 L  BOUND304
 L  BOUND404
 L  BOUND504
 L  BOUND604
 L  BOUND704
 L  BOUND804
 E  REGEN104
 E  REGEN204
 E  REGEN304
 E  REGEN404
 E  REGEN504
 E  REGEN604
 E  REGEN704
 E  REGEN804
 L  TFLOW104
 G  TFLOW204
 E  YIELD404
 L  BOUND305
 L  BOUND405
 L  BOUND505
 L  BOUND605
 L  BOUND705
 L  BOUND805
 E  REGEN105
 E  REGEN205
 E  REGEN305
 E  REGEN405
 E  REGEN505
 E  REGEN605
 E  REGEN705
 E  REGEN805
 L  TFLOW105
 G  TFLOW205
 E  YIELD505
 L  BOUND306
 L  BOUND406
 L  BOUND506
 L  BOUND606
 L  BOUND706
 L  BOUND806
 E  REGEN106
 E  REGEN206
 E  REGEN306
 E  REGEN406
 E  REGEN506
 E  REGEN606
 E  REGEN706
 E  REGEN806
 L  TFLOW106
 G  TFLOW206
 E  YIELD606
 L  BOUND307
 L  BOUND407
 L  BOUND507
 L  BOUND607
 L  BOUND707
 L  BOUND807
 E  REGEN107
 E  REGEN207
 E  REGEN307
 E  REGEN407
 E  REGEN507
 E  REGEN607
 E  REGEN707
 E  REGEN807
 L  TFLOW107
 G  TFLOW207
 E  YIELD707
COLUMNS
    CLASS301  BOUND301            1.   YIELD101           16.
    CLASS301  REGEN102       -.93742   REGEN402        .93742
    CLASS301  TFLOW102          14.4   TFLOW202          17.6
    CLASS401  BOUND401            1.   YIELD101          107.
    CLASS401  REGEN102       -.93742   REGEN502        .93742
    CLASS401  TFLOW102          96.3   TFLOW202         117.7
    CLASS501  BOUND501            1.   YIELD101          217.
    CLASS501  REGEN102       -.93742   REGEN602        .93742
    CLASS501  TFLOW102         195.3   TFLOW202         238.7
    CLASS601  BOUND601            1.   YIELD101          275.
    CLASS601  REGEN102       -.93742   REGEN702        .93742
    CLASS601  TFLOW102         247.5   TFLOW202         302.5
    CLASS701  BOUND701            1.   YIELD101          298.
    CLASS701  REGEN102       -.93742   REGEN802        .93742
    CLASS701  TFLOW102         268.2   TFLOW202         327.8
    CLASS801  BOUND801            1.   YIELD101          306.
    CLASS801  REGEN102       -.93742   REGEN802        .93742
    CLASS801  TFLOW102         275.4   TFLOW202         336.6
    STATE101  REGEN101            1.   REGEN102       -.06258
    STATE101  REGEN202       -.93742
    STATE201  REGEN201            1.   REGEN102       -.06258
    STATE201  REGEN302       -.93742
    STATE301  BOUND301           -1.   REGEN301            1.
    STATE301  REGEN102       -.06258   REGEN402       -.93742
    STATE401  BOUND401           -1.   REGEN401            1.
    STATE401  REGEN102       -.06258   REGEN502       -.93742
    STATE501  BOUND501           -1.   REGEN501            1.
    STATE501  REGEN102       -.06258   REGEN602       -.93742
    STATE601  BOUND601           -1.   REGEN601            1.
    STATE601  REGEN102       -.06258   REGEN702       -.93742
    STATE701  BOUND701           -1.   REGEN701            1.
    STATE701  REGEN102       -.06258   REGEN802       -.93742
    STATE801  BOUND801           -1.   REGEN801            1.
    STATE801  REGEN102       -.06258   REGEN802       -.93742
    BALAN101  HARV               -1.   YIELD101           -1.
    CLASS302  BOUND302            1.   YIELD202           16.
    CLASS302  REGEN103       -.93742   REGEN403        .93742
    CLASS302  TFLOW103          14.4   TFLOW203          17.6
    CLASS402  BOUND402            1.   YIELD202          107.
    CLASS402  REGEN103       -.93742   REGEN503        .93742
    CLASS402  TFLOW103          96.3   TFLOW203         117.7
    CLASS502  BOUND502            1.   YIELD202          217.
    CLASS502  REGEN103       -.93742   REGEN603        .93742
    CLASS502  TFLOW103         195.3   TFLOW203         238.7
    CLASS602  BOUND602            1.   YIELD202          275.
    CLASS602  REGEN103       -.93742   REGEN703        .93742
    CLASS602  TFLOW103         247.5   TFLOW203         302.5
    CLASS702  BOUND702            1.   YIELD202          298.
    CLASS702  REGEN103       -.93742   REGEN803        .93742
    CLASS702  TFLOW103         268.2   TFLOW203         327.8
    CLASS802  BOUND802            1.   YIELD202          306.
    CLASS802  REGEN103       -.93742   REGEN803        .93742
    CLASS802  TFLOW103         275.4   TFLOW203         336.6
    STATE102  REGEN102            1.   REGEN103       -.06258
    STATE102  REGEN203       -.93742
    STATE202  REGEN202            1.   REGEN103       -.06258
    STATE202  REGEN303       -.93742
    STATE302  BOUND302           -1.   REGEN302            1.
    STATE302  REGEN103       -.06258   REGEN403       -.93742
    STATE402  BOUND402           -1.   REGEN402            1.
    STATE402  REGEN103       -.06258   REGEN503       -.93742
    STATE502  BOUND502           -1.   REGEN502            1.
    STATE502  REGEN103       -.06258   REGEN603       -.93742
    STATE602  BOUND602           -1.   REGEN602            1.
    STATE602  REGEN103       -.06258   REGEN703       -.93742
    STATE702  BOUND702           -1.   REGEN702            1.
    STATE702  REGEN103       -.06258   REGEN803       -.93742
    STATE802  BOUND802           -1.   REGEN802            1.
    STATE802  REGEN103       -.06258   REGEN803       -.93742
    BALAN202  HARV          -.905063   TFLOW102           -1.
    BALAN202  TFLOW202           -1.   YIELD202           -1.
    PNLTY202  HARV               50.   TFLOW102           -1.
    CLASS303  BOUND303            1.   YIELD303           16.
    CLASS303  REGEN104       -.93742   REGEN404        .93742
    CLASS303  TFLOW104          14.4   TFLOW204          17.6
    CLASS403  BOUND403            1.   YIELD303          107.
    CLASS403  REGEN104       -.93742   REGEN504        .93742
    CLASS403  TFLOW104          96.3   TFLOW204         117.7
    CLASS503  BOUND503            1.   YIELD303          217.
    CLASS503  REGEN104       -.93742   REGEN604        .93742
    CLASS503  TFLOW104         195.3   TFLOW204         238.7
    CLASS603  BOUND603            1.   YIELD303          275.
    CLASS603  REGEN104       -.93742   REGEN704        .93742
    CLASS603  TFLOW104         247.5   TFLOW204         302.5
    CLASS703  BOUND703            1.   YIELD303          298.
    CLASS703  REGEN104       -.93742   REGEN804        .93742
    CLASS703  TFLOW104         268.2   TFLOW204         327.8
    CLASS803  BOUND803            1.   YIELD303          306.
    CLASS803  REGEN104       -.93742   REGEN804        .93742
    CLASS803  TFLOW104         275.4   TFLOW204         336.6
    STATE103  REGEN103            1.   REGEN104       -.06258
    STATE103  REGEN204       -.93742
    STATE203  REGEN203            1.   REGEN104       -.06258
    STATE203  REGEN304       -.93742
    STATE303  BOUND303           -1.   REGEN303            1.
    STATE303  REGEN104       -.06258   REGEN404       -.93742
    STATE403  BOUND403           -1.   REGEN403            1.
    STATE403  REGEN104       -.06258   REGEN504       -.93742
    STATE503  BOUND503           -1.   REGEN503            1.
    STATE503  REGEN104       -.06258   REGEN604       -.93742
    STATE603  BOUND603           -1.   REGEN603            1.
    STATE603  REGEN104       -.06258   REGEN704       -.93742
    STATE703  BOUND703           -1.   REGEN703            1.
    STATE703  REGEN104       -.06258   REGEN804       -.93742
    STATE803  BOUND803           -1.   REGEN803            1.
    STATE803  REGEN104       -.06258   REGEN804       -.93742
    BALAN303  HARV          -.819139   TFLOW103           -1.
    BALAN303  TFLOW203           -1.   YIELD303           -1.
    PNLTY303  HARV             45.25   TFLOW103           -1.
    CLASS304  BOUND304            1.   YIELD404           16.
    CLASS304  REGEN105       -.93742   REGEN405        .93742
    CLASS304  TFLOW105          14.4   TFLOW205          17.6
    CLASS404  BOUND404            1.   YIELD404          107.
    CLASS404  REGEN105       -.93742   REGEN505        .93742
    CLASS404  TFLOW105          96.3   TFLOW205         117.7
    CLASS504  BOUND504            1.   YIELD404          217.
    CLASS504  REGEN105       -.93742   REGEN605        .93742
    CLASS504  TFLOW105         195.3   TFLOW205         238.7
    CLASS604  BOUND604            1.   YIELD404          275.
    CLASS604  REGEN105       -.93742   REGEN705        .93742
    CLASS604  TFLOW105         247.5   TFLOW205         302.5
    CLASS704  BOUND704            1.   YIELD404          298.
    CLASS704  REGEN105       -.93742   REGEN805        .93742
    CLASS704  TFLOW105         268.2   TFLOW205         327.8
    CLASS804  BOUND804            1.   YIELD404          306.
    CLASS804  REGEN105       -.93742   REGEN805        .93742
    CLASS804  TFLOW105         275.4   TFLOW205         336.6
    STATE104  REGEN104            1.   REGEN105       -.06258
    STATE104  REGEN205       -.93742
    STATE204  REGEN204            1.   REGEN105       -.06258
    STATE204  REGEN305       -.93742
    STATE304  BOUND304           -1.   REGEN304            1.
    STATE304  REGEN105       -.06258   REGEN405       -.93742
    STATE404  BOUND404           -1.   REGEN404            1.
    STATE404  REGEN105       -.06258   REGEN505       -.93742
    STATE504  BOUND504           -1.   REGEN504            1.
    STATE504  REGEN105       -.06258   REGEN605       -.93742
    STATE604  BOUND604           -1.   REGEN604            1.
    STATE604  REGEN105       -.06258   REGEN705       -.93742
    STATE704  BOUND704           -1.   REGEN704            1.
    STATE704  REGEN105       -.06258   REGEN805       -.93742
    STATE804  BOUND804           -1.   REGEN804            1.
    STATE804  REGEN105       -.06258   REGEN805       -.93742
    BALAN404  HARV          -.741372   TFLOW104           -1.
    BALAN404  TFLOW204           -1.   YIELD404           -1.
    PNLTY404  HARV             40.96   TFLOW104           -1.
    CLASS305  BOUND305            1.   YIELD505           16.
    CLASS305  REGEN106       -.93742   REGEN406        .93742
    CLASS305  TFLOW106          14.4   TFLOW206          17.6
    CLASS405  BOUND405            1.   YIELD505          107.
    CLASS405  REGEN106       -.93742   REGEN506        .93742
    CLASS405  TFLOW106          96.3   TFLOW206         117.7
    CLASS505  BOUND505            1.   YIELD505          217.
    CLASS505  REGEN106       -.93742   REGEN606        .93742
    CLASS505  TFLOW106         195.3   TFLOW206         238.7
    CLASS605  BOUND605            1.   YIELD505          275.
    CLASS605  REGEN106       -.93742   REGEN706        .93742
    CLASS605  TFLOW106         247.5   TFLOW206         302.5
    CLASS705  BOUND705            1.   YIELD505          298.
    CLASS705  REGEN106       -.93742   REGEN806        .93742
    CLASS705  TFLOW106         268.2   TFLOW206         327.8
    CLASS805  BOUND805            1.   YIELD505          306.
    CLASS805  REGEN106       -.93742   REGEN806        .93742
    CLASS805  TFLOW106         275.4   TFLOW206         336.6
    STATE105  REGEN105            1.   REGEN106       -.06258
    STATE105  REGEN206       -.93742
    STATE205  REGEN205            1.   REGEN106       -.06258
    STATE205  REGEN306       -.93742
    STATE305  BOUND305           -1.   REGEN305            1.
    STATE305  REGEN106       -.06258   REGEN406       -.93742
    STATE405  BOUND405           -1.   REGEN405            1.
    STATE405  REGEN106       -.06258   REGEN506       -.93742
    STATE505  BOUND505           -1.   REGEN505            1.
    STATE505  REGEN106       -.06258   REGEN606       -.93742
    STATE605  BOUND605           -1.   REGEN605            1.
    STATE605  REGEN106       -.06258   REGEN706       -.93742
    STATE705  BOUND705           -1.   REGEN705            1.
    STATE705  REGEN106       -.06258   REGEN806       -.93742
    STATE805  BOUND805           -1.   REGEN805            1.
    STATE805  REGEN106       -.06258   REGEN806       -.93742
    BALAN505  HARV          -.670989   TFLOW105           -1.
    BALAN505  TFLOW205           -1.   YIELD505           -1.
    PNLTY505  HARV             37.07   TFLOW105           -1.
    CLASS306  BOUND306            1.   YIELD606           16.
    CLASS306  REGEN107       -.93742   REGEN407        .93742
    CLASS306  TFLOW107          14.4   TFLOW207          17.6
    CLASS406  BOUND406            1.   YIELD606          107.
    CLASS406  REGEN107       -.93742   REGEN507        .93742
    CLASS406  TFLOW107          96.3   TFLOW207         117.7
    CLASS506  BOUND506            1.   YIELD606          217.
    CLASS506  REGEN107       -.93742   REGEN607        .93742
    CLASS506  TFLOW107         195.3   TFLOW207         238.7
    CLASS606  BOUND606            1.   YIELD606          275.
    CLASS606  REGEN107       -.93742   REGEN707        .93742
    CLASS606  TFLOW107         247.5   TFLOW207         302.5
    CLASS706  BOUND706            1.   YIELD606          298.
    CLASS706  REGEN107       -.93742   REGEN807        .93742
    CLASS706  TFLOW107         268.2   TFLOW207         327.8
    CLASS806  BOUND806            1.   YIELD606          306.
    CLASS806  REGEN107       -.93742   REGEN807        .93742
    CLASS806  TFLOW107         275.4   TFLOW207         336.6
    STATE106  REGEN106            1.   REGEN107       -.06258
    STATE106  REGEN207       -.93742
    STATE206  REGEN206            1.   REGEN107       -.06258
    STATE206  REGEN307       -.93742
    STATE306  BOUND306           -1.   REGEN306            1.
    STATE306  REGEN107       -.06258   REGEN407       -.93742
    STATE406  BOUND406           -1.   REGEN406            1.
    STATE406  REGEN107       -.06258   REGEN507       -.93742
    STATE506  BOUND506           -1.   REGEN506            1.
    STATE506  REGEN107       -.06258   REGEN607       -.93742
    STATE606  BOUND606           -1.   REGEN606            1.
    STATE606  REGEN107       -.06258   REGEN707       -.93742
    STATE706  BOUND706           -1.   REGEN706            1.
    STATE706  REGEN107       -.06258   REGEN807       -.93742
    STATE806  BOUND806           -1.   REGEN806            1.
    STATE806  REGEN107       -.06258   REGEN807       -.93742
    BALAN606  HARV          -.607287   TFLOW106           -1.
    BALAN606  TFLOW206           -1.   YIELD606           -1.
    PNLTY606  HARV             33.55   TFLOW106           -1.
    CLASS307  HARV           198.203   BOUND307            1.
    CLASS307  YIELD707           16.
    CLASS407  HARV           222.975   BOUND407            1.
    CLASS407  YIELD707          107.
    CLASS507  HARV           252.173   BOUND507            1.
    CLASS507  YIELD707          217.
    CLASS607  HARV           281.025   BOUND607            1.
    CLASS607  YIELD707          275.
    CLASS707  HARV           292.467   BOUND707            1.
    CLASS707  YIELD707          298.
    CLASS807  HARV           296.446   BOUND807            1.
    CLASS807  YIELD707          306.
    STATE107  HARV          -159.355   REGEN107            1.
    STATE207  HARV          -177.186   REGEN207            1.
    STATE307  HARV          -198.203   BOUND307           -1.
    STATE307  REGEN307            1.
    STATE407  HARV          -222.975   BOUND407           -1.
    STATE407  REGEN407            1.
    STATE507  HARV          -252.173   BOUND507           -1.
    STATE507  REGEN507            1.
    STATE607  HARV          -281.025   BOUND607           -1.
    STATE607  REGEN607            1.
    STATE707  HARV          -292.467   BOUND707           -1.
    STATE707  REGEN707            1.
    STATE807  HARV          -296.446   BOUND807           -1.
    STATE807  REGEN807            1.
    BALAN707  HARV          -.549633   TFLOW107           -1.
    BALAN707  TFLOW207           -1.   YIELD707           -1.
    PNLTY707  HARV             30.36   TFLOW107           -1.
RHS
    RHS       REGEN101          .241   REGEN201          .125
    RHS       REGEN301         1.404   REGEN401         2.004
    RHS       REGEN501         9.768   REGEN601        16.385
    RHS       REGEN701         2.815   REGEN801        61.995
ENDATA
