"""Map server (continuous batching): a request's S(Q,w) is the same alone
and in a mix of concurrent requests of other trajectories, Q boxes, frame
counts, species sets and weights (the spread accumulates in fixed point and
every other stage is per request, so the maps agree bit for bit), and it
equals the S(Q,w) of the same request run through SqwEngine; X-ray weights
match a host evaluation of the form factors; admission respects the HBM
budget."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trajs():
    from mp_tools_amd import synth
    p1, c1 = synth.phonon_trajectory(4, 12, dt_ps=0.01, n_phonons=8, seed=11, device="cuda")
    p2, c2 = synth.amorphous_trajectory(18.0, 10, seed=12, device="cuda")
    return [dict(pos=p1, counts=c1, tpe=(2 * np.pi / 4,) * 3, qs=(2 * np.pi / (4 * synth.A_LATTICE),) * 3,
                 b=0.1 * synth.B_NEUTRON_FM, nsp=3, T=12),
            dict(pos=p2, counts=c2, tpe=(2 * np.pi / 18.0,) * 3, qs=(2 * np.pi / 18.0,) * 3,
                 b=0.1 * synth.B_AL_U_FM, nsp=2, T=10)]


def _server(trajs, **kw):
    from mp_tools_amd.serve import MapServer
    s = MapServer(**kw)
    for t in trajs:
        s.add_trajectory(t["pos"], t["counts"], t["tpe"], t["qs"])
    return s


def _run(server, reqs):
    from mp_tools_amd.serve import DONE
    tks = [server.submit(**r) for r in reqs]
    server.wait_idle()
    out = []
    for tk in tks:
        st, _, lat = server.poll(tk)
        assert st == DONE and lat > 0
        out.append(server.result(tk))
        server.release(tk)
    return out


REQS = [
    dict(traj=0, n_modes=(16, 16, 8), f0=2, T=8, mask=[1, 1, 1], n_f=4, weights=None),
    dict(traj=1, n_modes=(24, 16, 8), f0=0, T=10, mask=[1, 1], n_f=5, xray="seed"),
    dict(traj=0, n_modes=(32, 24, 16), f0=0, T=12, mask=[0, 1, 1], n_f=6, xray="seed"),
    dict(traj=1, n_modes=(16, 16, 16), f0=3, T=5, mask=[1, 0], n_f=2, weights=None),
]


def _materialise(trajs, reqs, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for r in reqs:
        r = dict(r)
        t = trajs[r["traj"]]
        if r.get("xray") == "seed":
            r["xray"] = np.concatenate([rng.uniform(0.5, 10, (t["nsp"], 4)), rng.uniform(0.5, 60, (t["nsp"], 4)),
                                        rng.uniform(0, 1, (t["nsp"], 1))], axis=1)
            r.pop("weights", None)
        else:
            r["weights"] = t["b"]
            r.pop("xray", None)
        out.append(r)
    return out


def test_alone_equals_mix_and_engine(gpu_lib, trajs):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    reqs = _materialise(trajs, REQS)
    alone = []
    for r in reqs:
        s = _server(trajs, max_active=1)
        alone.append(_run(s, [r])[0])
        s.close()
    mix = _run(_server(trajs, step_budget=1 << 20), reqs)   # tiny steps: frames interleave
    torch.cuda.synchronize()
    for k in range(len(reqs)):
        assert torch.equal(alone[k][0], mix[k][0]) and torch.equal(alone[k][1], mix[k][1]), f"request {k}"
    # request 0 through SqwEngine (neutron weights): same S(Q,w)
    r, t = reqs[0], trajs[0]
    eng = SqwEngine(3, r["n_modes"], t["counts"], eps=1e-6, weights=t["b"], tpe=t["tpe"], mask=r["mask"])
    amp = eng.amplitudes(t["pos"][r["f0"]:r["f0"] + r["T"]].contiguous())
    S, St = eng.reduce(amp, n_int=r["T"], dt=0.01, nat=float(np.sum(t["counts"])), n_f=r["n_f"])
    torch.cuda.synchronize()
    assert torch.equal(S, alone[0][0]) and torch.equal(St, alone[0][1])


def test_xray_weights_match_host_table(gpu_lib, trajs):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    r = _materialise(trajs, REQS)[1]
    t = trajs[1]
    S_srv, St_srv = _run(_server(trajs), [r])[0]
    N = r["n_modes"]
    ks = [np.arange(n) - n // 2 for n in N]
    K3, K2, K1 = np.meshgrid(ks[2], ks[1], ks[0], indexing="ij")
    q2 = ((K1 * t["qs"][0]) ** 2 + (K2 * t["qs"][1]) ** 2 + (K3 * t["qs"][2]) ** 2).ravel()
    x = r["xray"]
    tab = np.stack([x[s, 8] + sum(x[s, k] * np.exp(-x[s, 4 + k] * q2 / (16 * np.pi ** 2)) for k in range(4))
                    for s in range(t["nsp"])])
    eng = SqwEngine(3, N, t["counts"], eps=1e-6, weight_table=tab, tpe=t["tpe"])
    amp = eng.amplitudes(t["pos"][r["f0"]:r["f0"] + r["T"]].contiguous())
    S, St = eng.reduce(amp, n_int=r["T"], dt=0.01, nat=float(np.sum(t["counts"])), n_f=r["n_f"])
    torch.cuda.synchronize()
    # device FP64 form factors rounded to FP32 vs host FP64 rounded to FP32
    assert rel_l2(S_srv.cpu().numpy(), S.cpu().numpy()) < 1e-6
    assert rel_l2(St_srv.cpu().numpy(), St.cpu().numpy()) < 1e-6


def test_admission_budget(gpu_lib, trajs):
    reqs = _materialise(trajs, REQS)
    # budget for about one request at a time: all still complete, in order
    s = _server(trajs, hbm_budget=1 << 21)
    out = _run(s, reqs[:1] + reqs[3:])
    assert len(out) == 2
    from mp_tools_amd import _lib
    with pytest.raises(_lib.MpsqError):
        s.submit(traj=0, n_modes=(256, 256, 32), f0=0, T=12, weights=trajs[0]["b"])   # larger than the budget
