"""Writes crmath_tables.hpp: the double-double constants of the correctly rounded trigonometry (crmath.hpp).

Every value is computed with Python's decimal module at 80 significant digits (pi by Machin's formula,
sin / cos / atan by their Taylor series) and split into hi = RN(v), lo = RN(v - hi), so hi + lo carries
v to about 2^-107 relative. Run: python tomostar_amd/csrc/gen_crmath_tables.py
"""
from __future__ import annotations

import os
from decimal import Decimal, getcontext

getcontext().prec = 80


def arctan_inv(n: int) -> Decimal:
    """atan(1/n) for an integer n > 1 by its Taylor series."""
    x = Decimal(1) / n
    x2 = x * x
    term, total, k = x, x, 1
    eps = Decimal(10) ** -78
    while abs(term) > eps:
        term *= -x2
        k += 2
        total += term / k
    return total


PI = 16 * arctan_inv(5) - 4 * arctan_inv(239)  # Machin


def sin(x: Decimal) -> Decimal:
    term, total, k = x, x, 1
    while abs(term) > Decimal(10) ** -78:
        term *= -x * x / ((k + 1) * (k + 2))
        k += 2
        total += term
    return total


def cos(x: Decimal) -> Decimal:
    term, total, k = Decimal(1), Decimal(1), 0
    while abs(term) > Decimal(10) ** -78:
        term *= -x * x / ((k + 1) * (k + 2))
        k += 2
        total += term
    return total


def atan(x: Decimal) -> Decimal:
    """atan for 0 <= x <= 1: halve the argument three times (atan x = 2 atan(x / (1 + sqrt(1 + x^2)))),
    then the Taylor series."""
    halvings = 0
    while x > Decimal("0.13"):
        x = x / (1 + (1 + x * x).sqrt())
        halvings += 1
    term, total, k = x, x, 1
    while abs(term) > Decimal(10) ** -78:
        term *= -x * x
        k += 2
        total += term / k
    return total * (2 ** halvings)


def dd(v: Decimal) -> tuple[float, float]:
    hi = float(v)
    lo = float(v - Decimal(hi))
    return hi, lo


def split_bits(v: Decimal, bits: int) -> float:
    """v rounded to a double with only `bits` significant bits (so k * value is exact for small k)."""
    hi = float(v)
    m, e = __import__("math").frexp(hi)
    scale = 2.0 ** (bits - e)
    return round(Decimal(v) * Decimal(scale)) / scale


def h(x: float) -> str:
    return x.hex()


def main() -> None:
    out = []
    out.append("// Generated by gen_crmath_tables.py: double-double constants of crmath.hpp (hi + lo, ~2^-107).")
    out.append("#pragma once")
    out.append("")
    out.append("namespace ts {")
    out.append("namespace cr {")
    out.append("")
    pio2 = PI / 2
    p1 = split_bits(pio2, 33)
    r = pio2 - Decimal(p1)
    p2 = split_bits(r, 33)
    r -= Decimal(p2)
    p3 = float(r)
    r -= Decimal(p3)
    p4 = float(r)
    out.append("// pi/2 = kPio2_1 + kPio2_2 + kPio2_3 + kPio2_4 (33 + 33 + 53 + 53 bits): k * kPio2_1 and")
    out.append("// k * kPio2_2 are exact for |k| < 2^20 (Cody and Waite's reduction)")
    for name, v in (("kPio2_1", p1), ("kPio2_2", p2), ("kPio2_3", p3), ("kPio2_4", p4)):
        out.append(f"constexpr double {name} = {h(v)};")
    for name, v in (("kPio2", pio2), ("kPi", PI)):
        a, b = dd(v)
        out.append(f"constexpr double {name}_hi = {h(a)}, {name}_lo = {h(b)};")
    out.append(f"constexpr double kTwoOverPi = {h(float(2 / PI))};")
    out.append("")
    out.append("// Taylor coefficients as double-double where the error analysis needs them")
    coeffs = {
        "kS3": Decimal(-1) / 6, "kS5": Decimal(1) / 120, "kC4": Decimal(1) / 24, "kC6": Decimal(-1) / 720,
        "kA3": Decimal(-1) / 3, "kA5": Decimal(1) / 5, "kA7": Decimal(-1) / 7,
    }
    for name, v in coeffs.items():
        a, b = dd(v)
        out.append(f"constexpr double {name}_hi = {h(a)}, {name}_lo = {h(b)};")
    plain = {
        "kS7": Decimal(-1) / 5040, "kS9": Decimal(1) / 362880, "kS11": Decimal(-1) / 39916800,
        "kC8": Decimal(1) / 40320, "kC10": Decimal(-1) / 3628800, "kC12": Decimal(1) / 479001600,
        "kA9": Decimal(1) / 9, "kA11": Decimal(-1) / 11, "kA13": Decimal(1) / 13, "kA15": Decimal(-1) / 15,
    }
    for name, v in plain.items():
        out.append(f"constexpr double {name} = {h(float(v))};")
    out.append("")
    out.append("// sin(i/64) and cos(i/64), i = 0..51: {sin_hi, sin_lo, cos_hi, cos_lo}")
    out.append("constexpr double kSinCos64[52][4] = {")
    for i in range(52):
        a = Decimal(i) / 64
        s, c = dd(sin(a)), dd(cos(a))
        out.append(f"    {{{h(s[0])}, {h(s[1])}, {h(c[0])}, {h(c[1])}}},")
    out.append("};")
    out.append("")
    out.append("// atan(i/64), i = 0..64: {hi, lo}")
    out.append("constexpr double kAtan64[65][2] = {")
    for i in range(65):
        a = dd(atan(Decimal(i) / 64))
        out.append(f"    {{{h(a[0])}, {h(a[1])}}},")
    out.append("};")
    out.append("")
    out.append("}  // namespace cr")
    out.append("}  // namespace ts")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "crmath_tables.hpp")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
