"""The double-double constants of ne_atan2.hpp: atan(k/64) for k = 0..64,
1/(2m+1) for m = 1..10, pi and pi/2, each as (hi, lo) with hi = round(v) and
lo = round(v - hi), from an 80-digit evaluation (atan by halving the angle
until z <= 0.01, then its odd series).

    python nativeengine_amd/csrc/ne_atan2_tables.py   # prints the tables
"""
from decimal import Decimal, getcontext
getcontext().prec = 80
def datan(z):
    k = 0
    while z > Decimal("0.01"):
        z = z / (1 + (1 + z * z).sqrt()); k += 1
    s, t, n, z2 = Decimal(0), z, 1, z * z
    while True:
        term = t / n
        if abs(term) < Decimal(10) ** -75: break
        s += term if (n // 2) % 2 == 0 else -term
        t *= z2; n += 2
    return s * (2 ** k)
def dd(v):
    hi = float(v)
    lo = float(v - Decimal(hi))
    return hi.hex(), lo.hex()
PI = 4 * datan(Decimal(1))
print("PI", dd(PI)); print("PI_2", dd(PI / 2))
print("ATAN_TAB")
for k in range(65):
    h, l = dd(datan(Decimal(k) / 64))
    print(f"    {{{h}, {l}}},")
print("INV_ODD")
for m in range(1, 11):
    h, l = dd(Decimal(1) / (2 * m + 1))
    print(f"    {{{h}, {l}}},")
