NAME          RULES
* every MPS rule of the reference reader: second N row dropped, MARKER lines,
* duplicate entries, objective RHS constant, all RANGES cases, all BOUNDS types
ROWS
 N  COST
 N  OTHER
 L  LIM1
 G  LIM2
 E  MYEQN
 L  R4
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    X1        COST         1.0   LIM1         1.0
    X1        LIM2         1.0
    MARKER                 'MARKER'                 'INTEND'
    X2        COST         2.0   LIM1         1.0
    X2        MYEQN       -1.0   OTHER        5.0
    X3        COST        -1.0   MYEQN        1.0
    X3        R4           2.0   R4           3.0
    X4        LIM2         1.0
RHS
    RHS       COST        -7.5   LIM1         4.0
    RHS       LIM2         1.0   MYEQN        7.0
    RHS       R4           8.0
RANGES
    RNG       LIM1         2.5   LIM2         3.0
    RNG       MYEQN       -2.0
BOUNDS
 UP BND       X1           4.0
 MI BND       X2
 FX BND       X3           1.5
 FR BND       X4
 BV BND       NOPE
ENDATA
