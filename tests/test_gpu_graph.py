"""HIP-graph replay of the PME launch sequence (ops.PMEGraph).

Every replay, with new positions and charges copied into the static buffers
and eager work and synchronisations between replays, must equal an eager
evaluation of the same inputs: float64 to 1e-12 relative (float64 LDS atomics
leave last-bit differences between runs)."""

import numpy as np
import pytest
import torch

from torch_admp_amd import ops

pytestmark = pytest.mark.gpu

SYSTEMS = [(300, np.eye(3) * 16.0, [24, 24, 24]), (800, np.eye(3) * 24.0, [32, 30, 27]),
           (120, np.eye(3) * 11.0, [16, 18, 20])]


def _inputs(seed, dev):
    rng = np.random.default_rng(seed)
    pos, q = [], []
    for n, box, _ in SYSTEMS:
        pos.append(rng.random((n, 3)) @ box)
        c = rng.uniform(-1, 1, n)
        q.append(c - c.mean())
    return (torch.tensor(np.concatenate(pos), device=dev),
            torch.tensor(np.concatenate(q), device=dev))


@pytest.mark.parametrize("subset", [slice(0, 3), slice(1, 2)])
def test_graph_replays_equal_eager(dev, subset):
    systems = SYSTEMS[subset]
    plan = ops.PMEBatchPlan([s[0] for s in systems], [s[1] for s in systems], 5.5, 0.4,
                            [s[2] for s in systems], "float64")
    graph = ops.PMEGraph(plan, dev)
    sizes = [s[0] for s in systems]
    for it in range(4):
        pos_all, q_all = _inputs(10 + it, dev)
        keep = torch.cat([torch.arange(sum(s[0] for s in SYSTEMS[:k]),
                                       sum(s[0] for s in SYSTEMS[:k + 1]), device=dev)
                          for k in range(len(SYSTEMS))[subset]])
        pos, q = pos_all[keep], q_all[keep]
        e_g, f_g, g_g = (t.clone() for t in graph.run(pos, q))
        torch.cuda.synchronize()
        e_e, f_e, g_e = plan.run(pos, q)  # eager work between replays
        torch.cuda.synchronize()
        assert torch.allclose(e_g, e_e, rtol=1e-12, atol=1e-12 * float(e_e.abs().max())), it
        assert float((f_g - f_e).abs().max()) <= 1e-12 * max(1.0, float(f_e.abs().max())), it
        assert float((g_g - g_e).abs().max()) <= 1e-11 * max(1.0, float(g_e.abs().max())), it
        assert len(sizes) == e_g.shape[0]
