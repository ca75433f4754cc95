"""Netlist-level simulator of the generator's logical designs (test infrastructure).

It evaluates the *logical* design the generator wrote beside the program —
LUT4 truth tables, word ops on Python integers, SRAM arrays — without using
the serialized program, its nibble lookup tables, micropartitions or pools.
Agreement between this simulator and the oracle (which evaluates only the
serialized program) pins the oracle's evaluation of LUT, vector, memory,
exchange and register-copy semantics on every design the generator emits.

Register and memory images use the program's regPool / memPool positions so
results compare byte for byte.
"""
import numpy as np

K_REG, K_IN, K_LUT, K_CMP = 0, 1, 2, 3
W_ADD, W_SUB, W_EQ, W_LT, W_SHR, W_MUX, W_RDATA = range(7)
NONE = 0xFFFFFFFF
MASK64 = (1 << 64) - 1


def xorshift64(x):
    x ^= (x << 13) & MASK64
    x ^= x >> 7
    x ^= (x << 17) & MASK64
    return x


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


class Netlist:
    def __init__(self, blob: bytes):
        w = np.frombuffer(blob, dtype=np.uint32)
        assert w[0] == 0x4C4E4754 and w[1] == 1, "not a netlist file"
        i = 2

        def take(n, k):
            nonlocal i
            a = w[i:i + n * k].reshape(n, k) if k > 1 else w[i:i + n]
            i += n * k
            return a

        nb = int(w[i]); i += 1
        b = take(nb, 7).astype(np.int64)
        self.kind, self.a, self.b, self.c, self.d, self.tt, self.lvl = (b[:, j] for j in range(7))
        nr = int(w[i]); i += 1
        r = take(nr, 2).astype(np.int64)
        self.reg_pos, self.reg_next = r[:, 0], r[:, 1]
        ni = int(w[i]); i += 1
        self.in_pos = take(ni, 1).astype(np.int64)
        nwr = int(w[i]); i += 1
        self.wres = [tuple(int(x) for x in row) for row in take(nwr, 7)]
        nwg = int(w[i]); i += 1
        self.wregs = [tuple(int(x) for x in row) for row in take(nwg, 5)]
        ns = int(w[i]); i += 1
        self.srams = [tuple(int(x) for x in row) for row in take(ns, 13)]
        self.reg_pool = int(w[i]); self.mem_pool = int(w[i + 1]); i += 2
        assert i == len(w)
        self.luts = np.nonzero(self.kind == K_LUT)[0]
        self.levels = [self.luts[self.lvl[self.luts] == L] for L in range(1, int(self.lvl.max(initial=0)) + 1)]
        self.cmp_nodes = np.nonzero(self.kind == K_CMP)[0]


class NetlistSim:
    def __init__(self, net: Netlist, reg=None, mem=None):
        self.n = net
        self.reg = np.zeros(net.reg_pool, dtype=np.uint8) if reg is None else np.array(reg, dtype=np.uint8)
        self.mem = np.zeros(net.mem_pool, dtype=np.uint8) if mem is None else np.array(mem, dtype=np.uint8)
        self.xs = None
        self.cycles = 0

    def set_stimulus(self, seed, instance=0):
        s = splitmix64(seed + instance)
        self.xs = s if s else 0x9E3779B97F4A7C15

    def _word(self, pos, nib):
        v = 0
        for j in range(nib):
            v |= int(self.reg[pos + j]) << (4 * j)
        return v

    def step(self):
        n = self.n
        if self.xs is not None and len(n.in_pos):
            for d in range((len(n.in_pos) + 63) // 64):
                self.xs = xorshift64(self.xs)
                for j in range(64):
                    k = d * 64 + j
                    if k < len(n.in_pos):
                        self.reg[n.in_pos[k]] = (self.xs >> j) & 1
        val = np.zeros(len(n.kind), dtype=np.int64)
        m = n.kind == K_REG
        val[m] = self.reg[n.reg_pos[n.a[m]]]
        m = n.kind == K_IN
        val[m] = self.reg[n.in_pos[n.a[m]]]
        # word results, stage A (index order is topological)
        wval = [0] * len(n.wres)
        for r, (op, ak, a, bk, b, nib, sel) in enumerate(n.wres):
            mod = 1 << (4 * nib)
            if op == W_RDATA:
                depth, snib, k, ra_pos, *_rest = n.srams[a]
                base = n.srams[a][11]
                addr = self._word(ra_pos, k) & (depth - 1)
                v = 0
                for j in range(snib):
                    v |= int(self.mem[base + j * depth + addr]) << (4 * j)
                wval[r] = v
                continue
            va = wval[a] if ak else self._word(n.wregs[a][0], n.wregs[a][1])
            vb = wval[b] if bk else self._word(n.wregs[b][0], n.wregs[b][1])
            if op == W_ADD:
                wval[r] = (va + vb) % mod
            elif op == W_SUB:
                wval[r] = (va - vb) % mod
            elif op in (W_EQ, W_LT):
                sa = va - mod if va >> (4 * nib - 1) else va
                sb = vb - mod if vb >> (4 * nib - 1) else vb
                wval[r] = int(sa == sb) if op == W_EQ else int(sa < sb)
            elif op == W_SHR:
                s = vb & 15  # low nibble of the amount register
                wval[r] = va >> (4 * s)
            elif op == W_MUX:
                wval[r] = va if (val[sel] & 1) else vb
        for node in n.cmp_nodes:
            val[node] = wval[n.a[node]]
        # LUT4 levels: y = TT[d<<3 | a<<2 | b<<1 | c]
        for lv in n.levels:
            idx = ((val[n.d[lv]] & 1) << 3) | ((val[n.a[lv]] & 1) << 2) | ((val[n.b[lv]] & 1) << 1) | (val[n.c[lv]] & 1)
            val[lv] = (n.tt[lv] >> idx) & 1
        # next state (all from the current state)
        new_reg = self.reg.copy()
        new_reg[n.reg_pos] = val[n.reg_next]
        for pos, nib, sel, x, y in n.wregs:
            v = wval[x] if (sel == NONE or (val[sel] & 1)) else wval[y]
            for j in range(nib):
                new_reg[pos + j] = (v >> (4 * j)) & 15
        new_mem = self.mem.copy()
        for depth, snib, k, ra_pos, wa_pos, ra_mul, ra_add, wa_mul, wa_add, wdata, we, base, _rd in n.srams:
            mod = 1 << (4 * k)
            ra = self._word(ra_pos, k)
            wa = self._word(wa_pos, k)
            nra = (ra * (ra_mul % mod) + (ra_add % mod)) % mod
            nwa = (wa * (wa_mul % mod) + (wa_add % mod)) % mod
            for j in range(k):
                new_reg[ra_pos + j] = (nra >> (4 * j)) & 15
                new_reg[wa_pos + j] = (nwa >> (4 * j)) & 15
            if we == NONE or val[we] != 0:
                addr = wa & (depth - 1)
                v = wval[wdata]
                for j in range(snib):
                    new_mem[base + j * depth + addr] = (v >> (4 * j)) & 15
        self.reg = new_reg
        self.mem = new_mem
        self.cycles += 1

    def run(self, cycles):
        for _ in range(cycles):
            self.step()
