"""Host-side logic that needs no GPU: the benchmark's contract and the run
configurations."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

from felharness import REPO
from felmpf import config as C


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_metric_configuration_is_the_512x256x128_droplet():
    cfg = C.CONFIGS["droplet_impact"]()
    assert (cfg.ni, cfg.nj, cfg.nk) == (512, 128, 256)  # config (x,y,z) = (i,k,j)
    assert cfg.cells == 512 * 256 * 128
    assert abs(cfg.dx - cfg.dy) < 1e-15 and abs(cfg.dy - cfg.dz) < 1e-15  # cubic cells


def test_initial_state_layout_and_wall_maps():
    cfg = C.droplet_impact().scaled(128, 32, 64)
    st = C.initial_state(cfg)
    for name, a in st.items():
        assert a.shape == (cfg.nk + 6, cfg.nj + 6, cfg.ni + 6), name
    phi = st["phi1"]
    assert phi.min() >= 0.0 and phi.max() <= 1.0
    # the droplet (phi = 0, the liquid) exists and the ambient is phi = 1
    assert phi[3:-3, 3:-3, 3:-3].min() < 0.01 and phi[3:-3, 3:-3, 3:-3].max() > 0.99
    # wall maps live on the rows j = 1 and j = nj only
    for name in ("theta_0_array", "l1_array", "l2_array", "zeta_array"):
        m = st[name]
        rows = [j for j in range(m.shape[1]) if np.any(m[:, j, :] != 0.0)]
        assert rows == [3, cfg.nj + 2], (name, rows)
    # GNBC friction, srcs/main.f90:172: 0.21*6*(mu_l/l1 + mu_g/l2)
    z = st["zeta_array"][3, 3, 3]
    assert abs(z - 0.21 * 6.0 * (cfg.rmul / 0.02 + cfg.rmug / 0.02)) <= 1e-15 * z


def test_sor_pair_algorithmic_bytes():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(REPO, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    cfg = C.droplet_impact()
    # 4.75 elements of 8 bytes per interior cell
    assert abs(bench.sor_pair_bytes(cfg) - 8.0 * cfg.cells * 4.75) < 1.0


def test_observables_match_direct_sums():
    """The sensitivity module's observables against plain loops over cells."""
    torch = __import__("pytest").importorskip("torch")
    from felmpf import sensitivity as S
    cfg = C.patterned_wall().scaled(8, 4, 6)
    rng = np.random.default_rng(7)
    phi = rng.random((cfg.nk, cfg.nj, cfg.ni))
    got = S.observables(cfg, torch.from_numpy(phi)).numpy()
    liq, sx, sy, sz = 0.0, 0.0, 0.0, 0.0
    for k in range(cfg.nk):
        for j in range(cfg.nj):
            for i in range(cfg.ni):
                w = 1.0 - phi[k, j, i]
                liq += w
                sx += w * (i + 0.5) * cfg.dx
                sy += w * (j + 0.5) * cfg.dy
                sz += w * (k + 0.5) * cfg.dz
    area, ax, az = 0.0, 0.0, 0.0
    for k in range(cfg.nk):
        for i in range(cfg.ni):
            w = (1.0 - phi[k, 0, i]) * cfg.dx * cfg.dz
            area += w
            ax += w * (i + 0.5) * cfg.dx
            az += w * (k + 0.5) * cfg.dz
    ax, az = ax / area, az / area
    mom = sum((1.0 - phi[k, 0, i]) * cfg.dx * cfg.dz *
              (((i + 0.5) * cfg.dx - ax) ** 2 + ((k + 0.5) * cfg.dz - az) ** 2)
              for k in range(cfg.nk) for i in range(cfg.ni))
    want = np.array([sx / liq, sy / liq, sz / liq, area, mom])
    assert np.allclose(got, want, rtol=1e-13, atol=0.0)


def test_parameter_seeds_cover_the_bottom_wall_once():
    from felmpf import sensitivity as S
    cfg = C.patterned_wall().scaled(32, 16, 32)
    seeds = S.parameter_seeds(cfg)
    angles = [s["theta_0_array"] for s in seeds if "theta_0_array" in s]
    assert len(angles) == 2  # the two static angles of the checkerboard
    total = sum(angles)
    assert np.array_equal(total[:, 3, :], np.ones((cfg.nk + 6, cfg.ni + 6)))
    assert not np.any(np.delete(total, 3, axis=1))
    per_patch = S.patch_masks(cfg, per_patch=True)
    assert np.array_equal(sum(per_patch), total)
    slip = seeds[-1]
    ls = cfg.bottom.slip
    z = lambda s: C.zeta_of(cfg.rmul, cfg.rmug, s, s)  # noqa: E731
    dz = (z(ls * (1 + 1e-6)) - z(ls * (1 - 1e-6))) / (2e-6 * ls)
    assert abs(slip["zeta_array"][0, 3, 0] - dz) <= 1e-6 * abs(dz)
    assert np.array_equal(slip["l1_array"], slip["l2_array"])
