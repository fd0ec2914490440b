"""Program images for the sequencer: an encoder for the 64-bit instruction
words, the image container, the three ODE programs and their constant banks.

The image format is the reference's (host/src/program.c:68-77 and 1318-1431):
a 32-byte header - "CFTP", version 1, n_insns, n_consts, max_deposits, the
precision code, flags (BANK_EXT 1, SCRATCH_IO 2, SCRATCH_STRICT 4) and
scratch_io (slots in [15:0], out [31:16]) - then the constants (absent under
BANK_EXT), then the instruction words, all little-endian.

An instruction word (host/src/seq_exec.h:175-248): op [7:0], rd [11:8],
ra [15:12], rb [19:16], rc [23:20], rnd [26:24], ka/kb/kc [29:27], kx [30],
ctrl [31], imm [63:32].  A register's fifth bit lives in imm[24 + k] (k = 0
for rd, 1..3 for ra..rc); a constant index of 16 or more takes the kx form,
its low byte in imm[8k+7:8k] and its ninth bit in imm[28 + k].

The ODE programs restate the reference's generator (programs/gen_odes.py:
229-560): the instruction stream IS each program's definition, since it fixes
the order of every rounding, so each emitter below issues the same
operations in the same order on the same registers, which makes the image
the same bytes (checked against the reference's certificate hashes).
Lorenz-96 is generalised to any ring size K >= 4 with the same per-component
order; K = 40 is the reference's image.
"""

import struct
from fractions import Fraction

# opcodes (host/include/cft.h cft_op)
FMA, ADD, SUB, MUL, ABS, NEG = 0, 1, 2, 3, 4, 5
# control codes (host/src/seq_exec.h:63-72)
HALT, REPEAT, ENDREP, DEPOSIT, SETACT, ACTALL, STL, LDL, STX, LDX = range(10)
QUIET, ENDQUIET, RAISE = 12, 13, 14

FLAG_BANK_EXT, FLAG_SCRATCH_IO, FLAG_SCRATCH_STRICT = 1, 2, 4

# format code -> (width, precision p, exponent bits)
FORMATS = {0: (32, 24, 8), 1: (64, 53, 11), 2: (128, 113, 15), 3: (256, 237, 19)}


class Const(int):
    """A constant-bank index, as an ALU operand."""


def K(i):
    return Const(i)


def alu(op, rd, a=0, b=0, c=0, rnd=0, kx=None):
    """One ALU word.  a, b, c: register numbers, or Const bank indices."""
    srcs = (a, b, c)
    isk = [isinstance(s, Const) for s in srcs]
    if kx is None:
        kx = any(k and int(s) >= 16 for s, k in zip(srcs, isk))
    w = op | (rnd << 24) | (isk[0] << 27) | (isk[1] << 28) | (isk[2] << 29) | (int(kx) << 30)
    imm = 0
    w |= (rd & 0xF) << 8
    imm |= (rd >> 4) << 24
    for k, (s, is_k) in enumerate(zip(srcs, isk)):
        v = int(s)
        if is_k and kx:
            imm |= (v & 0xFF) << (8 * k)
            imm |= (v >> 8) << (28 + k)
        else:
            w |= (v & 0xF) << (12 + 4 * k)
            if not is_k:
                imm |= (v >> 4) << (25 + k)
    return w | (imm << 32)


def ctrl(code, rd=0, ra=0, rb=0, imm=0):
    w = code | (1 << 31) | ((rd & 0xF) << 8) | ((ra & 0xF) << 12) | ((rb & 0xF) << 16)
    imm |= ((rd >> 4) << 24) | ((ra >> 4) << 25) | ((rb >> 4) << 26)
    return w | (imm << 32)


def repeat(n):
    return ctrl(REPEAT, imm=n)


def endrep():
    return ctrl(ENDREP)


def halt():
    return ctrl(HALT)


def stl(ra, slot):
    return ctrl(STL, ra=ra, imm=slot)


def ldl(rd, slot):
    return ctrl(LDL, rd=rd, imm=slot)


def stx(ra, rb, step=0):
    return ctrl(STX, ra=ra, rb=rb, imm=step & 0xFFF)


def ldx(rd, rb, step=0):
    return ctrl(LDX, rd=rd, rb=rb, imm=step & 0xFFF)


def deposit(ra):
    return ctrl(DEPOSIT, ra=ra)


def setact(ra):
    return ctrl(SETACT, ra=ra)


def actall():
    return ctrl(ACTALL)


def quiet():
    return ctrl(QUIET)


def endquiet():
    return ctrl(ENDQUIET)


def raise_(ra):
    return ctrl(RAISE, ra=ra)


def image(fmt, insns, consts=(), n_consts=None, max_deposits=0, bank_ext=False,
          scratch_in=0, scratch_out=0, strict=False, scratch_io=None):
    """The image bytes.  consts: encodings (ints) stored in the image, unless
    bank_ext, where n_consts says how many the program addresses."""
    width = FORMATS[fmt][0]
    esz = width // 8
    flags = (FLAG_BANK_EXT if bank_ext else 0) | (FLAG_SCRATCH_STRICT if strict else 0)
    if scratch_io is None:
        scratch_io = scratch_in | (scratch_out << 16)
        if scratch_in or scratch_out:
            flags |= FLAG_SCRATCH_IO
    elif scratch_io:
        flags |= FLAG_SCRATCH_IO
    if n_consts is None:
        n_consts = len(consts)
    head = struct.pack("<4I4I", 0x50544643, 1, len(insns), n_consts, max_deposits, fmt,
                       flags, scratch_io)
    body = b"" if bank_ext else b"".join(int(v).to_bytes(esz, "little") for v in consts)
    return head + body + b"".join(struct.pack("<Q", w) for w in insns)


# ---- exact constants -------------------------------------------------------

def round_rational(fmt, q):
    """RN-even of the rational q (Fraction) in format fmt, as its encoding.
    Normal results only (the banks' constants are)."""
    width, p, ew = FORMATS[fmt]
    bias = (1 << (ew - 1)) - 1
    if q == 0:
        return 0
    sign = 1 if q < 0 else 0
    q = abs(q)
    num, den = q.numerator, q.denominator
    e = num.bit_length() - den.bit_length()
    if Fraction(num, den) < Fraction(2) ** e:
        e -= 1
    # 2^e <= q < 2^(e+1); significand scale 2^(p-1-e)
    s = p - 1 - e
    if s >= 0:
        m, r = divmod(num << s, den)
        rden = den
    else:
        m, r = divmod(num, den << -s)
        rden = den << -s
    if 2 * r > rden or (2 * r == rden and (m & 1)):
        m += 1
    if m >> p:
        m >>= 1
        e += 1
    be = e + bias
    assert 0 < be < (1 << ew) - 1, "normal results only"
    frac = m - (1 << (p - 1))
    return (sign << (width - 1)) | (be << (p - 1)) | frac


def to_rational(fmt, bits):
    """The value of a finite encoding, exactly."""
    width, p, ew = FORMATS[fmt]
    bias = (1 << (ew - 1)) - 1
    sign = bits >> (width - 1)
    be = (bits >> (p - 1)) & ((1 << ew) - 1)
    frac = bits & ((1 << (p - 1)) - 1)
    assert be != (1 << ew) - 1
    if be == 0:
        v = Fraction(frac) * Fraction(2) ** (1 - bias - (p - 1))
    else:
        v = Fraction(frac + (1 << (p - 1))) * Fraction(2) ** (be - bias - (p - 1))
    return -v if sign else v


def round_sqrt(fmt, q):
    """RN-even of sqrt(q) for a positive rational q (Fraction) in format fmt,
    as its encoding (normal results only).  An integer square root gives
    p + 2 bits of the root exactly, m, at scale 2^k; when the root is not
    exactly m * 2^k it lies strictly between m and m + 1, where every
    rounding decision is the same as for m + 1/2, so the rounding of that
    rational is the rounding of the root."""
    from math import isqrt
    _width, p, _ew = FORMATS[fmt]
    assert q > 0
    num, den = q.numerator, q.denominator
    e = (num.bit_length() - den.bit_length()) // 2 - 1
    while Fraction(2) ** (2 * (e + 1)) <= q:
        e += 1
    while Fraction(2) ** (2 * e) > q:
        e -= 1
    # 2^e <= sqrt(q) < 2^(e+1); m = floor(sqrt(q) * 2^(p+1-e)) has p+2 bits
    s = p + 1 - e
    scaled = q * Fraction(2) ** (2 * s)
    m = isqrt(scaled.numerator // scaled.denominator)
    exact = Fraction(m * m) == scaled
    v = Fraction(m) if exact else Fraction(2 * m + 1, 2)
    return round_rational(fmt, v / Fraction(2) ** s)


def dec(fmt, text):
    return round_rational(fmt, Fraction(text))


def bank_values(name, fmt, h="0.01"):
    """[(name, encoding)] in declaration order (programs/gen_odes.py:151-189):
    H = RN(h), H2 = H/2 (exact), H6 = RN(H/6), MH = -H, BETA = RN(8/3)."""
    H = dec(fmt, h)
    hq = to_rational(fmt, H)
    H2 = round_rational(fmt, hq / 2)
    H6 = round_rational(fmt, hq / 6)
    TWO = dec(fmt, "2")
    if name == "lorenz63-rk4":
        return [("H", H), ("H2", H2), ("H6", H6), ("TWO", TWO), ("SIGMA", dec(fmt, "10")),
                ("RHO", dec(fmt, "28")), ("BETA", round_rational(fmt, Fraction(8, 3)))]
    if name == "lorenz96-rk4":
        return [("H", H), ("H2", H2), ("H6", H6), ("TWO", TWO), ("F", dec(fmt, "8"))]
    if name == "henonheiles-lf":
        return [("H", H), ("H2", H2), ("MH", round_rational(fmt, -hq)), ("ONE", dec(fmt, "1")),
                ("TWO", TWO)]
    raise KeyError(name)


def bank_bytes(name, fmt, h="0.01"):
    esz = FORMATS[fmt][0] // 8
    return b"".join(v.to_bytes(esz, "little") for _n, v in bank_values(name, fmt, h))


# ---- the ODE programs --------------------------------------------------------

STEPS = {"lorenz63-rk4": 100, "lorenz96-rk4": 20, "henonheiles-lf": 100}


def lorenz63(fmt, steps=100):
    """Lorenz-63, classic RK4; state in scratch slots 0..2 (in and out)."""
    H, H2, H6, TWO, SIGMA, RHO, BETA = (K(i) for i in range(7))
    x, y, z, px, py, pz = 3, 4, 5, 6, 7, 8
    a = (9, 10, 11)
    b = (12, 13, 14)
    c = (15, 16, 17)
    t0, t1 = 18, 19
    P = []

    def f(ix, iy, iz, o):
        P.append(alu(SUB, t0, iy, 0, ix))           # y - x
        P.append(alu(MUL, o[0], SIGMA, t0))         # sigma (y - x)
        P.append(alu(SUB, t0, RHO, 0, iz))          # rho - z
        P.append(alu(NEG, t1, iy))
        P.append(alu(FMA, o[1], ix, t0, t1))        # x (rho - z) - y
        P.append(alu(MUL, t0, BETA, iz))            # beta z
        P.append(alu(NEG, t1, t0))
        P.append(alu(FMA, o[2], ix, iy, t1))        # x y - beta z

    st = (x, y, z)
    pp = (px, py, pz)
    P += [ldl(x, 0), ldl(y, 1), ldl(z, 2), repeat(steps)]
    f(x, y, z, a)
    P += [alu(FMA, pp[i], H2, a[i], st[i]) for i in range(3)]
    f(px, py, pz, b)
    P += [alu(FMA, c[i], TWO, b[i], a[i]) for i in range(3)]
    P += [alu(FMA, pp[i], H2, b[i], st[i]) for i in range(3)]
    f(px, py, pz, b)
    P += [alu(FMA, c[i], TWO, b[i], c[i]) for i in range(3)]
    P += [alu(FMA, pp[i], H, b[i], st[i]) for i in range(3)]
    f(px, py, pz, b)
    P += [alu(ADD, c[i], c[i], 0, b[i]) for i in range(3)]
    P += [alu(FMA, st[i], H6, c[i], st[i]) for i in range(3)]
    P += [endrep(), stl(x, 0), stl(y, 1), stl(z, 2), halt()]
    return image(fmt, P, n_consts=7, bank_ext=True, scratch_in=3, scratch_out=3)


def lorenz96(fmt, k=40, steps=20):
    """Lorenz-96 on a ring of k, classic RK4; every stage vector in scratch:
    Y at 0.., TA at k.., TB at 2k.., K1 at 3k.., the weighted sum at 4k..
    (5k slots); a four-register window slides round the ring."""
    if k < 4:
        raise ValueError("the window needs a ring of at least 4")
    H, H2, H6, TWO, F = (K(i) for i in range(5))
    W = (3, 4, 5, 6)
    t0, t1, kk, yy, aa = 7, 8, 9, 10, 11
    Y, TA, TB, K1, ACC = 0, k, 2 * k, 3 * k, 4 * k

    def w(j):
        return W[j % 4]

    P = [repeat(steps)]
    for s, IN in ((1, Y), (2, TB), (3, TA), (4, TB)):
        P += [ldl(w(-2), IN + k - 2), ldl(w(-1), IN + k - 1), ldl(w(0), IN), ldl(w(1), IN + 1)]
        for i in range(k):
            P.append(alu(SUB, t0, w(i + 1), 0, w(i - 2)))
            P.append(alu(SUB, t1, F, 0, w(i)))
            P.append(alu(FMA, kk, t0, w(i - 1), t1))
            if s == 1:
                P += [stl(kk, K1 + i), alu(FMA, t0, H2, kk, w(i)), stl(t0, TB + i)]
            elif s == 2:
                P += [ldl(yy, Y + i), alu(FMA, t0, H2, kk, yy), stl(t0, TA + i),
                      ldl(aa, K1 + i), alu(FMA, aa, TWO, kk, aa), stl(aa, ACC + i)]
            elif s == 3:
                P += [ldl(yy, Y + i), alu(FMA, t0, H, kk, yy), stl(t0, TB + i),
                      ldl(aa, ACC + i), alu(FMA, aa, TWO, kk, aa), stl(aa, ACC + i)]
            else:
                P += [ldl(aa, ACC + i), alu(ADD, aa, aa, 0, kk), ldl(yy, Y + i),
                      alu(FMA, yy, H6, aa, yy), stl(yy, Y + i)]
            if i + 2 <= k:
                P.append(ldl(w(i + 2), IN + (i + 2) % k))
    P += [endrep(), halt()]
    return image(fmt, P, n_consts=5, bank_ext=True, scratch_in=k, scratch_out=k)


def henonheiles(fmt, steps=100):
    """Henon-Heiles, Stormer-Verlet drift-kick-drift; x, y, px, py in
    scratch slots 0..3."""
    H, H2, MH, ONE, TWO = (K(i) for i in range(5))
    x, y, px, py, t0, t1, t2 = 3, 4, 5, 6, 7, 8, 9
    P = [ldl(x, 0), ldl(y, 1), ldl(px, 2), ldl(py, 3), repeat(steps)]
    P += [alu(FMA, x, H2, px, x), alu(FMA, y, H2, py, y),
          alu(FMA, t0, TWO, y, ONE), alu(MUL, t0, x, t0),
          alu(FMA, px, MH, t0, px),
          alu(MUL, t1, x, x), alu(NEG, t1, t1),
          alu(SUB, t2, y, 0, ONE), alu(FMA, t2, y, t2, t1),
          alu(FMA, py, H, t2, py),
          alu(FMA, x, H2, px, x), alu(FMA, y, H2, py, y)]
    P += [endrep(), stl(x, 0), stl(y, 1), stl(px, 2), stl(py, 3), halt()]
    return image(fmt, P, n_consts=5, bank_ext=True, scratch_in=4, scratch_out=4)


GENERATORS = {"lorenz63-rk4": lorenz63, "lorenz96-rk4": lorenz96,
              "henonheiles-lf": henonheiles}


def segment_trips(img):
    """The trip count of an image's first REPEAT (its segment length)."""
    n_insns, n_consts, _md, fmt, flags = struct.unpack_from("<5I", img, 8)[:5]
    esz = FORMATS[fmt][0] // 8
    off = 32 + (0 if flags & FLAG_BANK_EXT else n_consts * esz)
    for i in range(n_insns):
        w, = struct.unpack_from("<Q", img, off + 8 * i)
        if (w >> 31) & 1 and (w & 0xFF) == REPEAT:
            return w >> 32
    raise ValueError("no REPEAT in this image")


def with_trips(img, trips):
    """The same image with its first REPEAT's trip count replaced (how a
    segment of another length is built: one REPEAT, a different count)."""
    n_insns, n_consts, _md, fmt, flags = struct.unpack_from("<5I", img, 8)[:5]
    esz = FORMATS[fmt][0] // 8
    off = 32 + (0 if flags & FLAG_BANK_EXT else n_consts * esz)
    for i in range(n_insns):
        w, = struct.unpack_from("<Q", img, off + 8 * i)
        if (w >> 31) & 1 and (w & 0xFF) == REPEAT:
            w = (w & 0xFFFFFFFF) | (trips << 32)
            return img[:off + 8 * i] + struct.pack("<Q", w) + img[off + 8 * i + 8:]
    raise ValueError("no REPEAT in this image")
