NAME          CAPI
* docs/C_API.md LP: min x1 + x2 s.t. x1 + 2 x2 = 5, x2 <= 2, 3 x1 + 2 x2 <= 8, x free
ROWS
 N  obj
 E  r1
 L  r2
 L  r3
COLUMNS
    x1        obj          1.0   r1           1.0
    x1        r3           3.0
    x2        obj          1.0   r1           2.0
    x2        r2           1.0   r3           2.0
RHS
    rhs       r1           5.0   r2           2.0
    rhs       r3           8.0
BOUNDS
 FR bnd       x1
 FR bnd       x2
ENDATA
