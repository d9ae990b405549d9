NAME          MAXMIX
OBJSENSE
    MAX
ROWS
 N  PROFIT
 E  DEF
 G  FORCE
 L  CAP1
 L  CAP2
 E  LINK
COLUMNS
    F         PROFIT       1.0   DEF          2.0
    A         PROFIT       3.0   DEF          1.0
    A         CAP1         1.0   CAP2         2.0
    B         PROFIT       2.0   DEF         -1.0
    B         CAP1         1.0   LINK         1.0
    C         PROFIT       1.5   FORCE        1.0
    C         CAP2         1.0
    D         PROFIT      -1.0   FORCE        1.0
    D         CAP1         2.0
    E         PROFIT       4.0   LINK         2.0
    E         CAP2         1.0   CAP1         1.0
    G         PROFIT       0.5   CAP2         1.0
    G         CAP1         1.0   LINK         1.0
RHS
    RHS       DEF          4.0   FORCE        7.0
    RHS       CAP1        30.0   CAP2        25.0
    RHS       LINK         6.0
BOUNDS
 FR BND       F
 UP BND       A            6.0
 UP BND       B            6.0
 UP BND       C            3.0
 UP BND       D            4.0
 UP BND       E            5.0
 UP BND       G            5.0
ENDATA
