"""Hand-built edge-case programs on the GPU: every pool equal to the oracle and
to the closed-form answer, on both the plain and the wave64 kernels."""
import numpy as np
import pytest

import edge_programs as E
import oracle as O
from toucangpurt_amd import KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, POOL_EXCHANGE, POOL_MEM, POOL_REG, rt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", [KERNEL_PLAIN, KERNEL_WAVE64, KERNEL_STREAM], ids=["plain", "wave64", "stream"])
@pytest.mark.parametrize("case", E.all_cases(), ids=lambda c: c.name)
def test_edge_case_on_gpu(gpu_available, case, kernel):
    print(f"case {case.name} kernel={kernel}", flush=True)
    sim = rt.Simulator(0)
    assert sim.init(0, case.prog, None, False, case.want_log is not None) == 0
    sim.set_kernel(kernel)
    sim.write_pool(POOL_REG, case.reg)
    if case.mem.size:
        sim.write_pool(POOL_MEM, case.mem)
    stop = sim.eval_free_running(case.cycles)
    st = O.State(O.Program(case.prog), case.reg, case.mem)
    ostop = st.run(case.cycles, enable_print=case.want_log is not None)
    reg, mem, exg = sim.read_pool(POOL_REG), sim.read_pool(POOL_MEM), sim.read_pool(POOL_EXCHANGE)
    assert np.array_equal(reg, st.pool(0))
    assert np.array_equal(mem, st.pool(1))
    assert np.array_equal(exg, st.pool(2))
    for k, v in case.want_reg.items():
        assert reg[k] == v
    for k, v in case.want_mem.items():
        assert mem[k] == v
    if case.want_log is not None:
        assert sim.print_log() == case.want_log
    if case.want_stop is not None:
        assert stop == ostop == case.want_stop
        assert sim.stop_cycle() == st.stop_cycle
    sim.cleanup()
