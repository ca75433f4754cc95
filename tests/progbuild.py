"""Hand-built Toucan programs for tests (an independent writer of the format).

Writes GPUSimDesign.bin images in the order of src/Serializer.cpp:199-223
(partitions :147-166, microparts :106-121) with the F1 record layouts of
include/ToucanGPUGenDataTypes.h, using struct.pack with explicit zero
padding. It shares no code with the runtime's C++ writer.
"""
import struct

LUT_SIZE = 4642


def mid(lut, o0, o1, o2):
    """Packed middle-level op: lut<<18 | op0<<12 | op1<<6 | op2."""
    assert lut < (1 << 14) and max(o0, o1, o2) < 64
    return (lut << 18) | (o0 << 12) | (o1 << 6) | o2


class Part:
    def __init__(self, vp_size, consts=b""):
        self.vp_size = vp_size
        self.consts = bytes(consts)
        self.regread = []   # (reg, slot)
        self.exgread = []   # (exg, slot)
        self.levels = []    # [[mpart, ...], ...]
        self.regwrite = (0, 0, 0)   # (reg, dat, count)
        self.exgwrite = (0, 0, 0)   # (exg, dat, count)
        self.memwrite = []  # (multi, base, addrVec, dat, en)
        self.prints = []    # (en, msg)
        self.stops = []     # en

    def level(self, *mparts):
        self.levels.append(list(mparts))
        return self


def lut_mp(top, mids=(), last=()):
    """top: [(lut, op0, op1, op2)], mids: [[packed, ...]], last: [(shuffleId, slot)]."""
    return ("lut", list(top), [list(m) for m in mids], list(last))


def vecread_mp(ops):
    """ops: [(vecBase, vecLength, offset, isConst, i0, i1, i2, i3, outRange, result)]."""
    return ("vecread", list(ops))


def vecop_mp(ops):
    """ops: [(vec1Base, vec2Base, vecLength, opName, isV1V2Const, result)]."""
    return ("vecop", list(ops))


def memread_mp(ops):
    """ops: [(multi, memBase, en, addrVec, result)]."""
    return ("memread", list(ops))


class Prog:
    def __init__(self, reg_size, mem_size=0, exg_size=0, lut=None):
        self.reg_size, self.mem_size, self.exg_size = reg_size, mem_size, exg_size
        self.lut = bytes(lut) if lut is not None else bytes(LUT_SIZE)
        assert len(self.lut) == LUT_SIZE
        self.region = [[], []]
        self.parts = []
        self.msgs = []

    def add(self, region, part):
        self.parts.append((region, part))
        return part

    def serialize(self) -> bytes:
        order = [p for r in (0, 1) for (rr, p) in self.parts if rr == r]
        b = bytearray()
        u8 = lambda v: b.extend(struct.pack("<B", v))
        u32 = lambda v: b.extend(struct.pack("<I", v))
        u64 = lambda v: b.extend(struct.pack("<Q", v))

        def vec(raw_records, n):
            u64(n)
            b.extend(raw_records)

        u32(self.reg_size); u32(self.mem_size); u32(self.exg_size); u8(0)
        vec(self.lut, len(self.lut))
        vec(b"", 0)  # regPool (empty: Init fills)
        vec(b"", 0)  # memPool
        u64(len(order))
        for p in order:
            vec(b"", 0)  # valuePool
            u32(p.vp_size)
            vec(p.consts, len(p.consts))
            vec(b"".join(struct.pack("<IHxx", r, s) for r, s in p.regread), len(p.regread))
            vec(b"".join(struct.pack("<IHxx", e, s) for e, s in p.exgread), len(p.exgread))
            u64(len(p.levels))
            for lv in p.levels:
                u64(len(lv))
                for mp in lv:
                    kind = mp[0]
                    if kind == "lut":
                        _, top, mids, last = mp
                        u8(1)
                        vec(b"".join(struct.pack("<HHHH", *t) for t in top), len(top))
                        u64(len(mids))
                        for m in mids:
                            vec(b"".join(struct.pack("<I", w) for w in m), len(m))
                        vec(b"".join(struct.pack("<BxH", s, r) for s, r in last), len(last))
                    else:
                        u8(0)
                        ops = mp[1]
                        vr = ops if kind == "vecread" else []
                        vo = ops if kind == "vecop" else []
                        mr = ops if kind == "memread" else []
                        vec(b"".join(struct.pack("<HHHBxHHHHHH", *o) for o in vr), len(vr))
                        vec(b"".join(struct.pack("<HHHBBH", *o) for o in vo), len(vo))
                        vec(b"".join(struct.pack("<B7xQHHH2x", *o) for o in mr), len(mr))
            b.extend(struct.pack("<IHH", *p.regwrite))
            vec(b"".join(struct.pack("<B7xQHHH2x", *m) for m in p.memwrite), len(p.memwrite))
            vec(b"".join(struct.pack("<HH", *x) for x in p.prints), len(p.prints))
            vec(b"".join(struct.pack("<H", x) for x in p.stops), len(p.stops))
            b.extend(struct.pack("<IHH", *p.exgwrite))
        u64(2)
        ids = 0
        for r in (0, 1):
            n = sum(1 for rr, _ in self.parts if rr == r)
            vec(b"".join(struct.pack("<I", ids + k) for k in range(n)), n)
            ids += n
        u64(len(self.msgs))
        for m in self.msgs:
            vec(m.encode(), len(m.encode()))
        return bytes(b)


def symbols_blob(regs=(), signals=(), mems=()):
    """GPUSimIOSymbols.bin (src/Serializer.cpp:264-300). regs: [(name, [(id, width)])]."""
    b = bytearray()

    def s(x):
        e = x.encode()
        b.extend(struct.pack("<Q", len(e)))
        b.extend(e)

    for table, k in ((regs, 2), (signals, 3), (mems, 3)):
        b.extend(struct.pack("<Q", len(table)))
        for name, tuples in table:
            s(name)
            b.extend(struct.pack("<Q", len(tuples)))
            for t in tuples:
                assert len(t) == k
                b.extend(struct.pack("<" + "I" * k, *t))
    return bytes(b)


# Record sizes and field offsets of include/ToucanGPUGenDataTypes.h (x86-64).
RECORD_LAYOUT = {
    "CGRegReadMetaInfo": struct.calcsize("<IHxx"),
    "CGMicroPartLUTTopLevelOp": struct.calcsize("<HHHH"),
    "CGMicroPartLUTLastLevelWriteBack": struct.calcsize("<BxH"),
    "CGMicroPartVecRead": struct.calcsize("<HHHBxHHHHHH"),
    "CGMicroPartVecArithOrLogic": struct.calcsize("<HHHBBH"),
    "CGMicroPartMemRead": struct.calcsize("<B7xQHHH2x"),
    "CGMemWriteMetaInfo": struct.calcsize("<B7xQHHH2x"),
}
