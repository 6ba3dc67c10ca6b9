"""Graph replay of the step: the engine captures two steps (kernels only) per
field-pointer parity and replays them; the results must equal the eagerly
enqueued launches bit for bit on every replay, with eager work and host
synchronisations between the replays."""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state

pytestmark = pytest.mark.gpu

OBSTACLES = [(6.0, 12.0, 6.0, 12.0, 0.0, 5.0), (18.0, 26.0, 16.0, 24.0, 0.0, 8.0)]


def solver(cfg, case, fields, monkeypatch, graphs):
    monkeypatch.setenv("ELMM_GRAPH", "1" if graphs else "0")
    s = capi.Solver(cfg, case)
    cases.upload_initial(s, cfg, fields)
    s.init()
    return s


@pytest.mark.parametrize("precision,refine", [(capi.FP64, 0), (capi.FP32_REFINED, 0), (capi.FP32_REFINED, 1)])
def test_graph_replays_equal_eager(require_gpu, monkeypatch, precision, refine):
    cfg = capi.with_precision(cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05, nscalars=1), precision, refine)
    case = capi.Case(cfg, cuboids=OBSTACLES, sources=[(3.5, 3.5, 0.5, 1.0, 0)])
    f = cases.initial_fields(cfg, u0=3.0, amp=0.3, seed=9)
    eager = solver(cfg, case, f, monkeypatch, graphs=False)
    graph = solver(cfg, case, f, monkeypatch, graphs=True)
    # calls of both parities, odd and even lengths, sync and async forms;
    # every call after the first replays at least one captured graph
    for call, n in [("step", 1), ("step", 5), ("async", 4), ("step", 3), ("async", 5), ("async", 2), ("step", 6)]:
        for s in (eager, graph):
            if call == "step":
                s.step(n)
            else:
                s.step_async(n)
                s.wait()
        err = compare_state(graph, eager, cfg)
        assert max(err.values()) == 0.0, (call, n, err)
        # eager host work between replays: a divergence check on both
        assert graph.divergence() == eager.divergence()
    eager.close()
    graph.close()


def test_graph_replay_stretched_channel(require_gpu, monkeypatch):
    cfg = cases.channel_config(32, 16, 16)
    case = capi.Case(cfg)
    f = cases.channel_fields(cfg)
    eager = solver(cfg, case, f, monkeypatch, graphs=False)
    graph = solver(cfg, case, f, monkeypatch, graphs=True)
    for n in (3, 4, 5):
        eager.step(n)
        graph.step(n)
        err = compare_state(graph, eager, cfg, ["U", "V", "W", "Pr", "Visc"])
        assert max(err.values()) == 0.0, (n, err)
    eager.close()
    graph.close()
