"""Cycle-exact parity of the gfx950 kernels with the oracle (GPU).

Every case compares the full register, memory and exchange pools after every
cycle for a short trace (eval), then after a longer free run
(eval_free_running), on seeded designs from the generator.
"""
import pytest

from simharness import OracleRun, assert_pools_equal, gpu_pools, gpu_sim
from toucangpurt_amd import KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, gen

pytestmark = pytest.mark.gpu

KERNELS = [("plain", KERNEL_PLAIN), ("wave64", KERNEL_WAVE64), ("stream", KERNEL_STREAM)]

DESIGNS = {
    "lut20k": dict(preset="lut20k", seed=3),
    "mixed": dict(n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=600, n_srams=3, sram_depth=1024,
                  sram_width=32, part_luts=1200, word_part_ops=48, seed=5),
    "wide": dict(n_luts=8000, depth=12, n_regs=500, n_inputs=130, n_vecops=400, word_bits_mix=100, n_srams=2,
                 sram_depth=4096, sram_width=64, part_luts=700, cluster_luts=100, seed=9),
}

_cache = {}


def design(name):
    if name not in _cache:
        kw = dict(DESIGNS[name])
        _cache[name] = gen.generate(kw.pop("preset", None), **kw)
    return _cache[name]


@pytest.mark.parametrize("kname,kernel", KERNELS)
@pytest.mark.parametrize("dname", list(DESIGNS))
@pytest.mark.parametrize("randomize", [False, True])
def test_trace_then_free_run(gpu_available, kname, kernel, dname, randomize):
    d = design(dname)
    print(f"case {dname} kernel={kname} randomize={randomize}", flush=True)
    sim = gpu_sim(d.design, d.symbols, randomize, d.inputs, seed=17, kernel=kernel)
    orc = OracleRun(d.design, d.symbols, randomize, d.inputs, seed=17)
    assert_pools_equal(gpu_pools(sim), orc.pools(), "initial state")
    for c in range(12):
        sim.eval()
        orc.run(1)
        assert_pools_equal(gpu_pools(sim), orc.pools(), f"after cycle {c}")
    sim.eval_free_running(400)
    orc.run(400)
    assert_pools_equal(gpu_pools(sim), orc.pools(), "after 412 cycles")
    assert sim.cycles() == 412
    sim.cleanup()


@pytest.mark.parametrize("kname,kernel", KERNELS)
def test_print_and_stop(gpu_available, kname, kernel):
    d = gen.generate(None, n_luts=3000, depth=6, n_regs=100, n_inputs=20, n_prints=4, n_stops=2, part_luts=600,
                     cluster_luts=100, seed=21)
    print(f"case print/stop kernel={kname}", flush=True)
    sim = gpu_sim(d.design, d.symbols, False, d.inputs, seed=2, kernel=kernel, enable_print=True)
    orc = OracleRun(d.design, d.symbols, False, d.inputs, seed=2)
    stop_gpu = sim.eval_free_running(50)
    stop_orc = orc.run(50, enable_print=True)
    assert stop_gpu == stop_orc
    assert sim.stop_cycle() == orc.st.stop_cycle
    assert sim.print_log() == sorted(orc.st.print_log())
    assert_pools_equal(gpu_pools(sim), orc.pools(), "after 50 cycles")
    sim.cleanup()


@pytest.mark.parametrize("kname,kernel", KERNELS)
def test_instances_independent(gpu_available, kname, kernel):
    """Each instance of a context equals the oracle run with its own seed."""
    d = design("mixed")
    print(f"case instances kernel={kname}", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, None, kernel=kernel)
    sim.set_instances(3)
    sim.set_stimulus(d.inputs, 99)
    sim.run_batch([30, 7, 55])
    for k, cyc in enumerate([30, 7, 55]):
        orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=99, instance=k)
        orc.run(cyc)
        assert_pools_equal(gpu_pools(sim, k), orc.pools(), f"instance {k}")
    sim.cleanup()
