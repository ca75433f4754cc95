"""VCD waveform writer over snapshots (unsupport/SimVCD.cpp:99-241).

The variables are the symbol table's signals, ordered by (partition, value
slot) of their first fragment as SimVCD.cpp:105-116 orders them and read
through signal probes, followed by its registers in name order (state the
reference's writer cannot reach). Dotted names become nested scopes and
names without a dot go in scope "top" (SimVCD.cpp:99-176). Identifiers are
printable base-94 numbers from '!' (SimVCD.cpp:88-97). A dump writes
`#cycle` and the variables whose value changed since the previous dump
(SimVCD.cpp:193-241); multi-fragment values are the fragments' bits, most
significant fragment first.

Dumps are pipelined: dump(cycle) starts an asynchronous snapshot of the
register pool (device staging copy on the context stream, pinned host copy
on a side stream) and writes the previous snapshot, so the host copy and the
file writing overlap the next eval. close() writes the last one.
"""
import time

from . import _native


def vcd_identifier(i):
    s = ""
    while True:
        s = chr(ord("!") + i % 94) + s
        i //= 94
        if i == 0:
            return s


class VcdWriter:
    def __init__(self, sim, path, instance=0):
        self.sim = sim
        self.instance = instance
        self.f = open(path, "w")
        regs = sim.symbols["reg"] if sim.symbols else {}
        sigs = sim.symbols["signal"] if sim.symbols else {}
        sig_names = sorted((n for n in sigs if sigs[n]), key=lambda n: (sigs[n][0][0], sigs[n][0][1]))
        sim.probe_signals(sig_names)
        base = sim.pool_size(_native.POOL_REG)
        self.names = sig_names + sorted(regs)
        # per variable: (offset in the snapshot, width) per fragment
        self.frags = [[(base + sim._probe_first[n] + k, w) for k, (_, _, w) in enumerate(sigs[n])] for n in sig_names]
        self.frags += [[(rid, w) for rid, w in regs[n]] for n in sorted(regs)]
        self.width = [sum(w for _, w in fr) for fr in self.frags]
        self.ids = [vcd_identifier(i) for i in range(len(self.names))]
        self.last = [None] * len(self.names)
        self.pending = None  # (ticket, stamp)
        self._header()

    def _header(self):
        f = self.f
        f.write(f"$date\n    {time.strftime('%b %d %Y %H:%M:%S')}\n$end\n")
        f.write("$version\n    VCD generator\n$end\n$timescale\n    1ps\n$end\n")
        root = {"kids": {}, "sigs": []}
        for i, name in enumerate(self.names):
            node = root
            parts = name.split(".")
            if len(parts) == 1:
                node = node["kids"].setdefault("top", {"kids": {}, "sigs": []})
            for p in parts[:-1]:
                node = node["kids"].setdefault(p, {"kids": {}, "sigs": []})
            node["sigs"].append(i)

        def walk(name, node):
            f.write(f"$scope module {name} $end\n")
            for i in node["sigs"]:
                leaf = self.names[i].rsplit(".", 1)[-1]
                f.write(f"$var wire {self.width[i]} {self.ids[i]} {leaf} $end\n")
            for k in sorted(node["kids"]):
                walk(k, node["kids"][k])
            f.write("$upscope $end\n")

        for k in sorted(root["kids"]):
            walk(k, root["kids"][k])
        f.write("$enddefinitions $end\n$dumpvars\n")

    def _write(self, pool, stamp):
        out = [f"#{stamp}\n"]
        for i, fr in enumerate(self.frags):
            v = tuple(int(pool[rid]) for rid, _ in fr)
            if self.last[i] == v:
                continue
            self.last[i] = v
            if self.width[i] == 1:
                out.append(f"{v[0] & 1}{self.ids[i]}\n")
            else:
                bits = "".join(format(x & ((1 << w) - 1), f"0{w}b") for x, (_, w) in zip(v, fr))
                out.append(f"b{bits} {self.ids[i]}\n")
        self.f.write("".join(out))

    def dump(self, stamp):
        ticket = self.sim.snapshot_begin(self.instance)
        if self.pending is not None:
            self._flush_pending()
        self.pending = (ticket, stamp)

    def _flush_pending(self):
        ticket, stamp = self.pending
        pool, _ = self.sim.snapshot_wait(ticket)
        self._write(pool, stamp)
        self.pending = None

    def close(self):
        if self.f is None:
            return
        if self.pending is not None:
            self._flush_pending()
        self.f.close()
        self.f = None
