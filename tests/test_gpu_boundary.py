"""Boundary behaviour of the engine through the C ABI: where the reference's execute() throws
std::out_of_range for a StrRef field overflow (include/value_t.h:36-47, reached from copy_scan,
src/execute.cpp:323-358), the engine returns HJ_E_RANGE with the same message; NULL rows and the
last page that fits raise nothing."""
import pytest

import hjtest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pipeline", [1, 0])
@pytest.mark.parametrize("case", ["table", "column", "page"])
def test_strref_overflow_is_the_references_error(gpu_ctx, hjo, oracle, case, pipeline):
    plan, _ = hjtest.strref_plan(oracle.codec, case)
    with pytest.raises(RuntimeError) as want:
        oracle.execute(plan)
    msg = want.value.args[0].split(": ", 1)[1]
    gpu_ctx.set_option("pipeline", pipeline)
    try:
        with pytest.raises(hjo.EngineError) as e:
            gpu_ctx.execute(plan)
    finally:
        gpu_ctx.set_option("pipeline", 1)
    assert e.value.code == hjo.abi.HJ_E_RANGE
    assert e.value.msg == msg


@pytest.mark.parametrize("case", ["page_null", "none"])
def test_strref_within_limits(gpu_ctx, oracle, case):
    plan, _ = hjtest.strref_plan(oracle.codec, case)
    assert oracle.rows(gpu_ctx.execute(plan)) == oracle.rows(oracle.execute(plan))
