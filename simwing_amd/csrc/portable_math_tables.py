"""Generates the constant tables of portable_math.hpp with exact decimal
arithmetic (Python's decimal module at 80 significant digits).

Each double-double constant (hi, lo) has hi = the double nearest the real
value and lo = the double nearest the remainder. Printed as hexadecimal
floating literals; `python portable_math_tables.py` writes the block that
sits between the GENERATED markers of portable_math.hpp, and
`python portable_math_tables.py --check` verifies that the header holds it.
"""
from __future__ import annotations

import sys
from decimal import Decimal, getcontext
from pathlib import Path

getcontext().prec = 80
HEADER = Path(__file__).resolve().parent / "portable_math.hpp"
BEGIN = "// BEGIN GENERATED by portable_math_tables.py"
END = "// END GENERATED"


def pi() -> Decimal:
    """Machin: pi = 16 atan(1/5) - 4 atan(1/239)."""
    return 16 * atan_small(Decimal(1) / 5) - 4 * atan_small(Decimal(1) / 239)


def atan_small(x: Decimal) -> Decimal:
    """Maclaurin series, |x| <= 1/2."""
    total, term, n, x2 = Decimal(0), x, 0, x * x
    while True:
        piece = term / (2 * n + 1)
        if piece == 0 or abs(piece) < Decimal(10) ** -90:
            return total
        total += piece if n % 2 == 0 else -piece
        term *= x2
        n += 1


def atan(x: Decimal) -> Decimal:
    """atan(x) = 2 atan(x / (1 + sqrt(1 + x^2))) until |x| <= 1/2."""
    if abs(x) <= Decimal(1) / 2:
        return atan_small(x)
    return 2 * atan(x / (1 + (1 + x * x).sqrt()))


def sin_cos(x: Decimal) -> tuple[Decimal, Decimal]:
    s, c = Decimal(0), Decimal(0)
    term, n = Decimal(1), 0  # x^n / n!
    while abs(term) > Decimal(10) ** -90 or n < 4:
        if n % 2 == 0:
            c += term if (n // 2) % 2 == 0 else -term
        else:
            s += term if (n // 2) % 2 == 0 else -term
        n += 1
        term = term * x / n
    return s, c


def dd(value: Decimal) -> tuple[float, float]:
    hi = float(value)
    lo = float(value - Decimal(hi))
    return hi, lo


def truncate(value: float, bits: int) -> float:
    """value with its significand cut to `bits` leading bits."""
    import struct
    raw = struct.unpack("<Q", struct.pack("<d", value))[0]
    drop = 52 - (bits - 1)
    raw &= ~((1 << drop) - 1)
    return struct.unpack("<d", struct.pack("<Q", raw))[0]


def lit(x: float) -> str:
    return float.hex(x)


def block() -> str:
    p = pi()
    pio2 = p / 2
    lines = [BEGIN]
    hi, lo = dd(p)
    lines.append(f"constexpr double kPiHi = {lit(hi)}, kPiLo = {lit(lo)};")
    hi, lo = dd(pio2)
    lines.append(f"constexpr double kPio2Hi = {lit(hi)}, kPio2Lo = {lit(lo)};")
    lines.append(f"constexpr double kPio4 = {lit(float(p / 4))}, k3Pio4 = {lit(float(3 * p / 4))};")
    lines.append(f"constexpr double kTwoOverPi = {lit(float(2 / p))};")
    hi, lo = dd(Decimal(1) / 3)
    lines.append(f"constexpr double kThirdHi = {lit(hi)}, kThirdLo = {lit(lo)};")
    hi, lo = dd(Decimal(1) / 6)
    lines.append(f"constexpr double kSixthHi = {lit(hi)}, kSixthLo = {lit(lo)};")
    p1 = truncate(float(pio2), 33)
    p2 = truncate(float(pio2 - Decimal(p1)), 33)
    p3 = float(pio2 - Decimal(p1) - Decimal(p2))
    p3t = float(pio2 - Decimal(p1) - Decimal(p2) - Decimal(p3))
    lines.append("// pi/2 = kPio2Part1 + kPio2Part2 + kPio2Part3 + kPio2Part3Tail; parts 1 and 2 carry 33")
    lines.append("// significant bits, so n * part is exact for |n| < 2^20.")
    lines.append(f"constexpr double kPio2Part1 = {lit(p1)}, kPio2Part2 = {lit(p2)};")
    lines.append(f"constexpr double kPio2Part3 = {lit(p3)}, kPio2Part3Tail = {lit(p3t)};")
    lines.append("// atan(k / 64), k = 0..64, as (hi, lo).")
    lines.append("SWX_PM_TABLE double kAtanTable[65][2] = {")
    for k in range(65):
        hi, lo = dd(atan(Decimal(k) / 64))
        lines.append(f"    {{{lit(hi)}, {lit(lo)}}},")
    lines.append("};")
    lines.append("// sin(j / 64) and cos(j / 64), j = 0..51, as (sin hi, sin lo, cos hi, cos lo).")
    lines.append("SWX_PM_TABLE double kSinCosTable[52][4] = {")
    for j in range(52):
        s, c = sin_cos(Decimal(j) / 64)
        sh, sl = dd(s)
        ch, cl = dd(c)
        lines.append(f"    {{{lit(sh)}, {lit(sl)}, {lit(ch)}, {lit(cl)}}},")
    lines.append("};")
    lines.append(END)
    return "\n".join(lines) + "\n"


def header_block() -> str:
    text = HEADER.read_text()
    start = text.index(BEGIN)
    stop = text.index(END) + len(END) + 1
    return text[start:stop]


if __name__ == "__main__":
    if "--check" in sys.argv:
        ok = header_block() == block()
        print("tables match" if ok else "tables differ")
        sys.exit(0 if ok else 1)
    sys.stdout.write(block())
