NAME          FORCING
ROWS
 N  COST
 G  FLOW
 L  CAP
 L  MIX
 L  BAL
COLUMNS
    X1        COST         1.0   FLOW         1.0
    X1        CAP          1.0   MIX          1.0
    X2        COST         2.0   FLOW         1.0
    X2        BAL          1.0
    X3        COST        -1.0   MIX         -1.0
    X3        CAP          1.0
    X4        COST         3.0   MIX          1.0
    X4        BAL          2.0   CAP          1.0
RHS
    RHS       FLOW         4.0   CAP         10.0
    RHS       MIX         -3.0   BAL          8.0
BOUNDS
 UP BND       X1           2.0
 UP BND       X2           2.0
 UP BND       X3           5.0
 UP BND       X4           9.0
ENDATA
