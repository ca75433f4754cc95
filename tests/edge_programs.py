"""Hand-built programs that pin each evaluation rule, with closed-form answers.

Each builder returns a Case: the serialized program, initial pools, a cycle
count and the expected register / memory pools (or a checker), computed here
from the rule's statement in the reference (file:line in each docstring),
not by running any evaluator. The table below is deliberately asymmetric so
that swapped operands, '+' for '|', a wrong immediate threshold or a wrong
lane modulus give a different answer.
"""
from dataclasses import dataclass, field

import numpy as np

from progbuild import LUT_SIZE, Part, Prog, lut_mp, memread_mp, mid, vecop_mp, vecread_mp

TABLE = bytes([0] + [((i * 37) ^ (i >> 5) ^ (i >> 9)) & 15 for i in range(1, LUT_SIZE)])
L1, L2 = 100, 500


def T(pos):
    """Table read with the reference's u16 index; past the table reads 0 (defined here)."""
    pos &= 0xFFFF
    return TABLE[pos] if pos < LUT_SIZE else 0


@dataclass
class Case:
    name: str
    prog: bytes
    reg: np.ndarray
    mem: np.ndarray
    cycles: int
    want_reg: dict = field(default_factory=dict)   # reg index -> value
    want_mem: dict = field(default_factory=dict)
    want_log: list = None
    want_stop: bool = None


def _pools(reg_size, mem_size, regs=None, mem=None):
    r = np.zeros(reg_size, np.uint8)
    m = np.zeros(mem_size, np.uint8)
    for k, v in (regs or {}).items():
        r[k] = v
    for k, v in (mem or {}).items():
        m[k] = v
    return r, m


def top_operands():
    """Top level (SimEval.cu:336-360): < 16 is an immediate, else a slot; index 0 passes op2."""
    regs = [5, 9, 3, 12, 7, 1, 14, 2]
    p = Prog(64, lut=TABLE)
    P = p.add(0, Part(128))
    P.regread = [(k, 20 + k) for k in range(8)]
    P.level(lut_mp(
        top=[(L1, 20, 21, 22), (L1, 3, 21, 22), (0, 20, 21, 23), (0, 0, 0, 15), (L2, 15, 15, 24)],
        last=[(32, 64), (33, 65), (34, 66), (35, 67), (36, 68), (0, 69), (0, 70), (0, 71)]))
    P.regwrite = (16, 64, 8)
    r, m = _pools(64, 0, dict(enumerate(regs)))
    want = [T(L1 + (5 << 8 | 9 << 4 | 3)), T(L1 + (3 << 8 | 9 << 4 | 3)), 12, 15, T(L2 + (15 << 8 | 15 << 4 | 7)),
            0, 0, 0]
    return Case("top_operands", p.serialize(), r, m, 1, {16 + k: v for k, v in enumerate(want)})


def middle_levels():
    """Middle levels (SimEval.cu:365-420): operands < 32 are immediates, 32..63 a lane of
    the previous level (mod 32); lanes past a level's size hold 0. Last level (:422-447):
    shuffleId < 32 is an immediate, else lane shuffleId mod 32."""
    regs = [5, 9, 3, 12]
    p = Prog(64, lut=TABLE)
    P = p.add(0, Part(128))
    P.regread = [(k, 20 + k) for k in range(4)]
    m1 = [mid(L1, 32, 33, 34), mid(0, 0, 0, 35), mid(L1, 3, 33, 20), mid(0, 0, 0, 63), mid(0, 0, 0, 31)]
    m2 = [mid(L2, 32, 33, 34), mid(0, 0, 0, 35), mid(0, 0, 0, 36), mid(0, 0, 0, 33)]
    P.level(lut_mp(top=[(0, 0, 0, 20), (0, 0, 0, 21), (0, 0, 0, 22), (0, 0, 0, 23)], mids=[m1, m2],
                   last=[(32, 64), (33, 65), (34, 66), (35, 67), (200, 68), (7, 69), (0, 70), (0, 71)]))
    P.regwrite = (16, 64, 8)
    prev = [5, 9, 3, 12] + [0] * 28
    a0 = T(L1 + (prev[0] << 8 | prev[1] << 4 | prev[2]))
    a2 = T(L1 + (3 << 8 | (prev[1] << 4) | 20))  # op2 = immediate 20, ORed (not added)
    lvl1 = [a0, prev[3], a2, prev[31], 31] + [0] * 27
    b0 = T(L2 + (lvl1[0] << 8 | lvl1[1] << 4 | lvl1[2]))
    lvl2 = [b0, lvl1[3], lvl1[4], lvl1[1]] + [0] * 28
    want = [lvl2[0], lvl2[1], lvl2[2], lvl2[3], lvl2[200 % 32], 7, 0, 0]
    r, m = _pools(64, 0, dict(enumerate(regs)))
    return Case("middle_levels", p.serialize(), r, m, 1, {16 + k: v for k, v in enumerate(want)})


def last_fanout():
    """Last level (SimEval.cu:422-447) sending one lane's value to several slots, next to an
    immediate and a lane past the level's size; a second micropart of the level shares the
    wave (lanes 32-63)."""
    regs = [5, 9, 3, 12]
    p = Prog(64, lut=TABLE)
    P = p.add(0, Part(128))
    P.regread = [(k, 20 + k) for k in range(4)]
    P.level(lut_mp(top=[(L1, 20, 21, 22), (0, 0, 0, 23)],
                   last=[(32, 64), (32, 65), (33, 66), (32, 67), (9, 68), (40, 69)]),
            lut_mp(top=[(0, 0, 0, 21)], last=[(32, 70), (32, 71)]))
    P.regwrite = (16, 64, 8)
    a0 = T(L1 + (5 << 8 | 9 << 4 | 3))
    want = [a0, a0, 12, a0, 9, 0, 9, 9]
    r, m = _pools(64, 0, dict(enumerate(regs)))
    return Case("last_fanout", p.serialize(), r, m, 1, {16 + k: v for k, v in enumerate(want)})


def vector_read():
    """VecRead (SimEval.cu:449-481): off = (i0<<12|i1<<8|i2<<4|i3) + offset; inside the
    vector read vp/const, else outRange (< 16 immediate)."""
    regs = [5, 9, 3, 12, 7, 1, 14, 2]
    consts = bytes([10, 11, 12, 13, 14, 15])
    p = Prog(64, lut=TABLE)
    P = p.add(0, Part(128, consts))
    P.regread = [(k, 20 + k) for k in range(8)]
    P.level(vecread_mp([
        (20, 8, 2, 0, 0, 0, 0, 21, 7, 64),   # off = 9 + 2 = 11 >= 8 -> immediate 7
        (20, 8, 1, 0, 0, 0, 0, 22, 7, 65),   # off = 3 + 1 = 4 -> vp[24] = 7
        (0, 6, 0, 1, 0, 0, 0, 25, 23, 66),   # const, off = 1 -> 11
        (0, 6, 5, 1, 0, 0, 0, 2, 23, 67),    # off = 7 >= 6 -> slot 23 = 12
        (20, 8, 0, 0, 1, 0, 0, 0, 7, 68),    # i0 immediate 1 -> off 4096 -> 7
        (20, 8, 0, 0, 0, 0, 0, 27, 0, 69),   # off = 2 -> vp[22] = 3
    ]), lut_mp(top=[], last=[(0, 70), (0, 71)]))
    P.regwrite = (16, 64, 8)
    r, m = _pools(64, 0, dict(enumerate(regs)))
    want = [7, 7, 11, 12, 7, 3, 0, 0]
    return Case("vector_read", p.serialize(), r, m, 1, {16 + k: v for k, v in enumerate(want)})


def _nibbles(v, n):
    return [(v >> (4 * k)) & 15 for k in range(n)]


def _word(nibs):
    return sum(int(x) << (4 * k) for k, x in enumerate(nibs))


def vector_ops():
    """VecArithOrLogic (SimEval.cu:482-548): little-endian nibbles, both operands sign
    extended from bit 4*len-1 (sign_extend_128, :222-228); ADD/SUB/MUL write len nibbles
    (wrapping), EQ/LT/LE compare signed and write 0/1, unknown ops write 0."""
    A, B = 0x8001, 0x7FFF                   # -32767, 32767 as 16-bit
    W1 = (1 << 127) | 0x0123456789ABCDEF    # 128-bit operands
    W2 = (1 << 127) - 5
    C = 0x00F0
    regs = {}
    for k, v in enumerate(_nibbles(A, 4) + _nibbles(B, 4)):
        regs[k] = v
    for k, v in enumerate(_nibbles(W1, 32) + _nibbles(W2, 32)):
        regs[8 + k] = v
    consts = bytes(_nibbles(C, 4))
    p = Prog(256, lut=TABLE)
    P = p.add(0, Part(512, consts))
    P.regread = [(k, 16 + k) for k in range(8)] + [(8 + k, 100 + k) for k in range(64)]
    ops = [
        (16, 20, 4, 0, 0, 200), (16, 20, 4, 1, 0, 204), (16, 20, 4, 2, 0, 208),
        (16, 20, 4, 4, 0, 212), (16, 20, 4, 5, 0, 213), (16, 20, 4, 3, 0, 214), (16, 20, 4, 9, 0, 215),
        (20, 16, 4, 4, 0, 216),                  # 32767 < -32767 -> 0
        (16, 0, 4, 0, 1, 220),                   # v2 from constVecPool
        (100, 132, 32, 0, 0, 224), (100, 132, 32, 4, 0, 256), (100, 132, 32, 1, 0, 260),
    ]
    P.level(vecop_mp(ops[:9]), vecop_mp(ops[9:]))
    P.level(lut_mp(top=[], last=[(0, 217), (0, 218), (0, 219), (0, 257), (0, 258), (0, 259)]))
    P.regwrite = (128, 200, 92)
    M16, M128 = 1 << 16, 1 << 128

    def s(v, bits):
        return v - (1 << bits) if v >> (bits - 1) else v

    want = {}
    res = {200: _nibbles((A + B) % M16, 4), 204: _nibbles((A - B) % M16, 4), 208: _nibbles((A * B) % M16, 4),
           212: [int(s(A, 16) < s(B, 16))], 213: [int(s(A, 16) <= s(B, 16))], 214: [int(A == B)], 215: [0],
           216: [int(s(B, 16) < s(A, 16))], 217: [0, 0, 0], 220: _nibbles((A + C) % M16, 4),
           224: _nibbles((W1 + W2) % M128, 32), 256: [int(s(W1, 128) < s(W2, 128))], 257: [0, 0, 0],
           260: _nibbles((W1 - W2) % M128, 32)}
    for slot, nibs in res.items():
        for k, v in enumerate(nibs):
            want[128 + slot - 200 + k] = v
    r, m = _pools(256, 0, regs)
    return Case("vector_ops", p.serialize(), r, m, 1, want)


def _vec_value(nibs, n):
    """Operand of SimEval.cu:491-503 as the reference forms it: OR of byte k << 4k over
    __int128, then sign_extend_128 from bit 4n-1 when 0 < n < 32 (:222-228)."""
    v = 0
    for k in range(n):
        v |= int(nibs[k]) << (4 * k)
    v &= (1 << 128) - 1
    if 0 < n < 32 and (v >> (4 * n - 1)) & 1:
        v |= ((1 << 128) - 1) & ~((1 << (4 * n)) - 1)
    return v - (1 << 128) if v >> 127 else v


def vector_ops_narrow():
    """VecArithOrLogic with operands of at most 16 nibbles (the ops of one level share a
    wave): sign extension from bits 63, 31 and 3, compares that depend on bits above 63
    (a top byte above 15 ORed in, as the reference's byte OR does), length 0 (value 0, no
    nibbles written), an unknown op. Expected values follow the reference's __int128
    arithmetic (SimEval.cu:482-548) restated here."""
    ops_vals = {
        "xa": _nibbles(0x8000000000000001, 16), "ya": _nibbles(0x7FFFFFFFFFFFFFFF, 16),
        "xb": _nibbles(0x3123456789ABCDEF, 15) + [0x23], "yb": _nibbles(0x3123456789ABCDEF, 16),
        "x8": _nibbles(0x12345678, 7) + [0x31], "y8": _nibbles(0x12345678, 8),
        "n8": [0x8], "n7": [0x7],
    }
    regs, base, slot = {}, {}, 0
    for name, nibs in ops_vals.items():
        base[name] = 16 + slot
        for k, v in enumerate(nibs):
            regs[slot + k] = v
        slot += len(nibs)
    ADD, SUB, MUL, EQ, LT, LE = 0, 1, 2, 3, 4, 5
    spec = [("xa", "ya", 16, ADD), ("xa", "ya", 16, SUB), ("xa", "ya", 16, MUL), ("xa", "ya", 16, LT),
            ("xa", "ya", 16, LE), ("xa", "ya", 16, EQ), ("ya", "xa", 16, LT), ("xb", "yb", 16, EQ),
            ("yb", "xb", 16, LT), ("xb", "yb", 16, LT), ("xb", "yb", 16, LE), ("xb", "yb", 16, SUB),
            ("x8", "y8", 8, EQ), ("y8", "x8", 8, LT), ("x8", "y8", 8, ADD), ("x8", "y8", 8, MUL),
            ("n8", "n7", 1, LT), ("n8", "n7", 1, ADD), ("xa", "ya", 0, EQ), ("xa", "ya", 0, LT),
            ("xa", "ya", 16, 9)]
    while (sum(n if op <= MUL else 1 for _, _, n, op in spec)) % 4:  # whole words for the register copy
        spec.append(("xb", "xb", 16, EQ))
    ops, want_slots, res = [], {}, 200
    for a, b, n, op in spec:
        ops.append((base[a], base[b], n, op, 0, res))
        va, vb = _vec_value(ops_vals[a], n), _vec_value(ops_vals[b], n)
        if op <= MUL:
            r = (va + vb if op == ADD else va - vb if op == SUB else va * vb) % (1 << 128)
            want_slots[res] = [(r >> (4 * k)) & 15 for k in range(n)]
            res += n
        else:
            want_slots[res] = [int(va == vb) if op == EQ else int(va < vb) if op == LT
                               else int(va <= vb) if op == LE else 0]
            res += 1
    ops.append((base["xa"], base["ya"], 0, ADD, 0, 500))  # length 0: writes nothing
    count = res - 200
    p = Prog(512, lut=TABLE)
    P = p.add(0, Part(640))
    P.regread = [(k, 16 + k) for k in range(slot)]
    P.level(vecop_mp(ops))
    P.regwrite = (256, 200, count)
    want = {}
    for sl, nibs in want_slots.items():
        for k, v in enumerate(nibs):
            want[256 + sl - 200 + k] = v
    r, m = _pools(512, 0, regs)
    return Case("vector_ops_narrow", p.serialize(), r, m, 1, want)


def memory():
    """MemRead (SimEval.cu:550-575): disabled reads write nothing; address from 8 slots,
    x4 when hasMultipleWriter. MemWrite (:670-693): same address rule, dat < 16 immediate."""
    mem0 = {i: (i * 5 + 1) & 15 for i in range(64)}
    regs = {0: 3, 1: 1, 8: 2}
    p = Prog(64, 64, lut=TABLE)
    P = p.add(0, Part(128))
    P.regread = [(k, 20 + k) for k in range(8)] + [(8 + k, 28 + k) for k in range(8)] + [(16, 36)]
    P.level(lut_mp(top=[(0, 0, 0, 9)], last=[(32, 66), (0, 69), (0, 70), (0, 71)]))
    P.level(memread_mp([(0, 0, 1, 20, 64), (1, 2, 1, 28, 65), (0, 0, 36, 20, 66), (0, 1, 1, 20, 67),
                        (0, 50, 1, 20, 68)]))
    P.memwrite = [(0, 40, 20, 6, 1), (0, 0, 28, 64, 36), (1, 1, 28, 20, 1)]
    P.regwrite = (16, 64, 8)
    r, m = _pools(64, 64, regs, mem0)
    want_reg = {16: mem0[19], 17: mem0[2 + 8], 18: 9, 19: mem0[20], 20: 0, 21: 0, 22: 0, 23: 0}
    # base 50 + 19 = 69 is past the 64-nibble pool: reads 0 (defined here)
    want_mem = dict(mem0)
    want_mem[59] = 6
    want_mem[1 + 8] = 3  # multi-writer: (2 << 2) + 1, data = slot 20 = 3
    return Case("memory", p.serialize(), r, m, 1, want_reg, want_mem)


def word_copies():
    """Last level (SimEval.cu:625-761): ceil(count/4) 4-nibble words from both ends aligned
    down; when exchangeWrite.count != 0 only the exchange copy runs (:847-849)."""
    regs = {k: (k * 3 + 1) & 15 for k in range(8)}
    p = Prog(64, 0, 16, lut=TABLE)
    P0 = p.add(0, Part(64))
    P0.regread = [(k, 20 + k) for k in range(8)]
    P0.exgwrite = (5, 21, 5)      # words vp[20..28) -> exg[4..12)
    P0.regwrite = (0, 20, 4)      # ignored
    P0.level(lut_mp(top=[], last=[(0, 40)]))
    P1 = p.add(1, Part(64))
    P1.exgread = [(4 + k, 40 + k) for k in range(8)]
    P1.regwrite = (17, 41, 6)     # words vp[40..48) -> reg[16..24)
    r, m = _pools(64, 0, regs)
    want = {16 + k: regs[k] for k in range(8)}
    want.update({k: regs[k] for k in range(8)})
    return Case("word_copies", p.serialize(), r, m, 1, want)


def print_stop(fire):
    """Print (SimEval.cu:696-704) when en != 0; stop (:706-722) sets the sticky flag."""
    p = Prog(16, lut=TABLE)
    p.msgs = ["a\n", "b\n"]
    P = p.add(0, Part(64))
    P.regread = [(0, 20), (1, 21)]
    P.level(lut_mp(top=[], last=[(0, 30)]))
    P.prints = [(20, 0), (21, 1), (1, 1), (0, 0)]
    P.stops = [21, 20 if fire else 0]
    r, m = _pools(16, 0, {0: 1, 1: 0})
    log = [(c, msg) for c in range(3) for msg in (0, 1)]
    return Case(f"print_stop_{fire}", p.serialize(), r, m, 3, {0: 1, 1: 0}, want_log=log, want_stop=fire)


def counter64():
    """Known answer: a 64-bit register incremented by a constant each cycle equals N mod 2^64."""
    p = Prog(16, lut=TABLE)
    P = p.add(0, Part(64, bytes(_nibbles(1, 16))))
    P.regread = [(k, 16 + k) for k in range(16)]
    P.level(vecop_mp([(16, 0, 16, 0, 1, 32)]))
    P.regwrite = (0, 32, 16)
    r, m = _pools(16, 0, {k: 15 for k in range(16)})  # start at 2^64 - 1
    n = 300
    want = {k: v for k, v in enumerate(_nibbles((2 ** 64 - 1 + n) % 2 ** 64, 16))}
    return Case("counter64", p.serialize(), r, m, n, want)


XOR = 2835  # table region of o1 ^ o2 (o0 = 0) in the tables below


def xor_table():
    t = bytearray(TABLE)
    for i in range(256):
        t[XOR + i] = (i >> 4) ^ (i & 15)
    return bytes(t)


def _xor_tree(slots):
    """One 32-lane LUT micropart computing the XOR of the given slots (<= 32 terms)."""
    top = []
    terms = list(slots)
    if len(terms) % 2:
        terms.append(0)  # immediate 0
    for k in range(0, len(terms), 2):
        top.append((XOR, 0, terms[k], terms[k + 1]))
    mids = []
    n = len(top)
    while n > 1:
        lvl = []
        for k in range(0, n, 2):
            lvl.append(mid(XOR, 0, 32 + k, 32 + k + 1) if k + 1 < n else mid(0, 0, 0, 32 + k))
        mids.append(lvl)
        n = len(lvl)
    return top, mids


def gf2_xorshift_rows():
    """Rows of the GF(2) matrix of xorshift64 (Marsaglia 2003, shifts 13, 7, 17)."""
    cols = []
    for j in range(64):
        x = 1 << j
        x ^= (x << 13) & (2 ** 64 - 1)
        x ^= x >> 7
        x ^= (x << 17) & (2 ** 64 - 1)
        cols.append(x)
    return [[j for j in range(64) if (cols[j] >> i) & 1] for i in range(64)]


def xorshift_in_logic(seed=0x9E3779B97F4A7C15, n=40):
    """Known answer: 64 one-bit registers holding the xorshift64 state, next state built from
    XOR trees of nibble LUT ops; after n cycles they hold xorshift64^n(seed)."""
    rows = gf2_xorshift_rows()
    p = Prog(128, 0, 64, lut=xor_table())
    P0 = p.add(0, Part(256))
    P0.regread = [(k, 16 + k) for k in range(64)]
    mps = []
    for i, row in enumerate(rows):
        top, mids = _xor_tree([16 + j for j in row])
        mps.append(lut_mp(top, mids, [(32, 128 + i)]))
    P0.level(*mps)
    P0.exgwrite = (0, 128, 64)
    P1 = p.add(1, Part(128))
    P1.exgread = [(k, 16 + k) for k in range(64)]
    P1.regwrite = (0, 16, 64)
    x = seed
    for _ in range(n):
        x ^= (x << 13) & (2 ** 64 - 1)
        x ^= x >> 7
        x ^= (x << 17) & (2 ** 64 - 1)
    r, m = _pools(128, 0, {k: (seed >> k) & 1 for k in range(64)})
    return Case("xorshift_in_logic", p.serialize(), r, m, n, {k: (x >> k) & 1 for k in range(64)})


def lfsr16(n=100, seed=0xACE1):
    """Known answer: 16-bit Fibonacci LFSR x^16 + x^14 + x^13 + x^11 + 1 (bit i = register i,
    feedback into bit 0, shift toward bit 15)."""
    p = Prog(32, 0, 16, lut=xor_table())
    P0 = p.add(0, Part(128))
    P0.regread = [(k, 16 + k) for k in range(16)]
    top, mids = _xor_tree([16 + 15, 16 + 13, 16 + 12, 16 + 10])
    # next state: bit 0 = feedback, bit i = bit i-1
    last = [(32, 48)]
    P0.level(lut_mp(top, mids, last))
    P0.level(lut_mp(top=[(0, 0, 0, 48)] + [(0, 0, 0, 16 + i - 1) for i in range(1, 16)],
                    last=[(32 + i, 64 + i) for i in range(16)]))
    P0.exgwrite = (0, 64, 16)
    P1 = p.add(1, Part(64))
    P1.exgread = [(k, 16 + k) for k in range(16)]
    P1.regwrite = (0, 16, 16)
    s = seed
    for _ in range(n):
        fb = ((s >> 15) ^ (s >> 13) ^ (s >> 12) ^ (s >> 10)) & 1
        s = ((s << 1) | fb) & 0xFFFF
    r, m = _pools(32, 0, {k: (seed >> k) & 1 for k in range(16)})
    return Case("lfsr16", p.serialize(), r, m, n, {k: (s >> k) & 1 for k in range(16)})


def all_cases():
    return [top_operands(), middle_levels(), last_fanout(), vector_read(), vector_ops(), vector_ops_narrow(), memory(),
            word_copies(),
            print_stop(False), print_stop(True), counter64(), xorshift_in_logic(), lfsr16()]
