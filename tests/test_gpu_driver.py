"""The C++ driver (softx-d-21-00092_amd/main_optimized, the reference's
main_optimized structure over the C ABI) end to end on the GPU: MV initial
condition, Langevin steps, measurement and the reference's output file
(field.h:2521-2547), for both engines, against the oracle fed the same Philox
stream.  The output file prints 7 significant digits, which bounds the check."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "softx-d-21-00092_amd", "main_optimized")


def write_cfg(path, N, stat, mass, Ny, steps, meas, step, coupling, engine, seed=11, precision="fp64", extra=""):
    path.write_text(f"""stat = {stat}
mu = 30.72
mass = {mass}
elementaryWilsonLines = {Ny}
file_name = out_{engine}_{precision}
langevin_steps = {steps}
measurements = {meas}
step = {step}
EvolutionChoice = MOMENTUM_EVOLUTION
CouplingChoice = {coupling}
KernelChoice = SIN_KERNEL
N = {N}
seed = {seed}
engine = {engine}
precision = {precision}
""" + extra)


def read_out(path):
    rows = [l.split() for l in open(path) if l.strip()]
    return np.array([[float(v) for v in r] for r in rows])


def oracle_output(orc, jw, N, stat, mass, Ny, steps, meas, step, coupling, seed=11, trajs=None, position=False,
                  kspace=False):
    """main_optimized.cpp:159-355 through the oracle with the driver's Philox stream
    (stat sample s has the noise key {seed, trajs[s]}, trajs[s] = s by default);
    kspace: the position-space normals of the momentum-space stream."""
    mu = float(np.float32(30.72)) / N
    m = float(np.float32(mass))
    ds = float(np.float32(step))
    every = max(steps // meas, 1)
    V = N * N
    sums = [np.zeros(V) for _ in range(meas)]
    errs = [np.zeros(V) for _ in range(meas)]
    for s in (trajs if trajs is not None else range(stat)):
        z = np.concatenate([orc.philox_normals(N, seed, s, k, jw.TAG_MV, 8) for k in range(Ny)])
        U = orc.mv_init(N, z, Ny, 1.0, mu, m)        # Poisson mass m g^2 mu inside
        for l in range(steps):
            zl = orc.kspace_normals(N, seed, s, Ny + l, jw.TAG_KSPACE) if kspace else \
                orc.philox_normals(N, seed, s, Ny + l, jw.TAG_LANGEVIN, 16)
            if position:           # main_optimized.cpp:243-300, POSITION_EVOLUTION
                xx, xy = orc.generate_gaussian(N, zl)
                if coupling == 1:
                    xx, xy = orc.correlate_noise(xx, N), orc.correlate_noise(xy, N)
                A, B = orc.prepare_A_and_B_local(xx, xy, U, N, coupling, 1)
                U = orc.update_uf(U, B, A, ds, N)
            else:
                U = orc.langevin_step(U, zl, N, coupling, 1, ds)
            if l % every == 0:             # main_optimized.cpp:317-336
                t = l * meas // steps
                c = orc.measure_corr(U, N)[:, 0].real
                sums[t] += c
                errs[t] += c * c
    ph, _, _ = orc.momenta(N)
    x = 1.0 / 3.0 / stat
    rows = []
    for xx in range(N):
        for yy in range(N):
            if abs(xx - yy) <= 4:
                i = xx * N + yy
                # sqrt(phat2) as the file prints it ("%f", field.h:2543)
                rows.append([meas - 1, xx, yy, float("%f" % np.sqrt(ph[i])), x * ph[i] * sums[-1][i],
                             x * ph[i] * errs[-1][i]])
    return np.array(rows)


@pytest.mark.parametrize("coupling,code", [("NOISE_COUPLING_CONSTANT", 1), ("FIXED_COUPLING", 3)])
def test_driver_end_to_end(jw, orc, tmp_path, coupling, code):
    assert os.path.exists(DRIVER), "driver not built"
    N, stat, mass, Ny, steps, meas, step = 32, 2, 2.0, 6, 8, 4, 0.0004
    outs = {}
    for engine in ("fused", "reference"):
        cfg = tmp_path / f"cfg_{engine}"
        write_cfg(cfg, N, stat, mass, Ny, steps, meas, step, coupling, engine)
        r = subprocess.run([DRIVER, "1", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "Evolution time" in r.stdout
        outs[engine] = read_out(tmp_path / f"out_{engine}_fp64_{N}_{N}_mpi1_r0.dat")
    ref = oracle_output(orc, jw, N, stat, mass, Ny, steps, meas, step, code)
    for engine, got in outs.items():
        assert got.shape == ref.shape
        assert np.array_equal(got[:, :3], ref[:, :3])
        np.testing.assert_allclose(got[:, 3:], ref[:, 3:], rtol=2e-6, atol=1e-12)


@pytest.mark.parametrize("engine,precision,rtol", [("batched", "fp64", 2e-6), ("fused", "mixed", 1e-4),
                                                   ("batched", "mixed", 1e-4)])
def test_driver_batched_and_mixed(jw, orc, tmp_path, engine, precision, rtol):
    """The batched engine (all stat samples in one batch) and the mixed-precision
    form, end to end against the oracle on the same Philox stream; FP32 fields
    are compared at 1e-4 of the printed values."""
    assert os.path.exists(DRIVER), "driver not built"
    N, stat, mass, Ny, steps, meas, step = 64, 3, 2.0, 6, 8, 4, 0.0004
    cfg = tmp_path / "cfg"
    write_cfg(cfg, N, stat, mass, Ny, steps, meas, step, "FIXED_COUPLING", engine, precision=precision)
    r = subprocess.run([DRIVER, "1", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Evolution time" in r.stdout
    got = read_out(tmp_path / f"out_{engine}_{precision}_{N}_{N}_mpi1_r0.dat")
    ref = oracle_output(orc, jw, N, stat, mass, Ny, steps, meas, step, 3)
    assert got.shape == ref.shape and np.array_equal(got[:, :3], ref[:, :3])
    np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=0, atol=5e-7)     # printed with %f
    np.testing.assert_allclose(got[:, 4:], ref[:, 4:], rtol=rtol, atol=1e-12)


@pytest.mark.parametrize("engine,precision,rtol", [("fused", "fp64", 2e-6), ("batched", "fp64", 2e-6),
                                                   ("fused", "mixed", 1e-4)])
def test_driver_momentum_noise(jw, orc, tmp_path, engine, precision, rtol):
    """noise_space = momentum: the fused and batched engines draw the noise spectra
    per mode; the oracle steps on the position-space normals of the same stream."""
    assert os.path.exists(DRIVER), "driver not built"
    N, stat, mass, Ny, steps, meas, step = 64, 2, 2.0, 6, 8, 4, 0.0004
    cfg = tmp_path / "cfg"
    write_cfg(cfg, N, stat, mass, Ny, steps, meas, step, "FIXED_COUPLING", engine, precision=precision,
              extra="noise_space = momentum\n")
    r = subprocess.run([DRIVER, "1", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = read_out(tmp_path / f"out_{engine}_{precision}_{N}_{N}_mpi1_r0.dat")
    ref = oracle_output(orc, jw, N, stat, mass, Ny, steps, meas, step, 3, kspace=True)
    pos = oracle_output(orc, jw, N, stat, mass, Ny, steps, meas, step, 3)
    assert got.shape == ref.shape and np.array_equal(got[:, :3], ref[:, :3])
    np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=0, atol=5e-7)     # printed with %f
    np.testing.assert_allclose(got[:, 4:], ref[:, 4:], rtol=rtol, atol=1e-12)
    assert not np.allclose(pos[:, 4:], ref[:, 4:], rtol=10 * rtol, atol=1e-12)     # another noise


def test_driver_rejects_momentum_noise_in_reference_engine(tmp_path):
    assert os.path.exists(DRIVER), "driver not built"
    cfg = tmp_path / "cfg"
    write_cfg(cfg, 64, 1, 2.0, 2, 2, 1, 0.0004, "FIXED_COUPLING", "reference", extra="noise_space = momentum\n")
    r = subprocess.run([DRIVER, "1", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0


def test_driver_measurement_schedule_uneven(jw, orc, tmp_path):
    """langevin_steps = 10, measurements = 3: the reference measures after steps 0, 3, 6
    and 9 into slots 0, 0, 1, 2 and prints slot 2, the state after step 9
    (main_optimized.cpp:317-336, 353)."""
    assert os.path.exists(DRIVER), "driver not built"
    N, stat, mass, Ny, steps, meas, step = 32, 1, 2.0, 4, 10, 3, 0.0004
    cfg = tmp_path / "cfg"
    write_cfg(cfg, N, stat, mass, Ny, steps, meas, step, "FIXED_COUPLING", "fused")
    r = subprocess.run([DRIVER, "1", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = read_out(tmp_path / f"out_fused_fp64_{N}_{N}_mpi1_r0.dat")
    ref = oracle_output(orc, jw, N, stat, mass, Ny, steps, meas, step, 3)
    assert got.shape == ref.shape and np.array_equal(got[:, :3], ref[:, :3])
    np.testing.assert_allclose(got[:, 3:], ref[:, 3:], rtol=2e-6, atol=1e-12)


@pytest.mark.parametrize("engine,precision", [("fused", "fp64"), ("reference", "fp64"), ("fused", "mixed")])
def test_driver_checkpoint_resume_bit_exact(jw, tmp_path, engine, precision):
    """A run stopped after some steps (stop_after) and resumed from its checkpoint
    (resume = 1) writes the same output file, byte for byte, as the run that was
    not stopped: U, the Philox state and the measurement accumulators resume."""
    assert os.path.exists(DRIVER), "driver not built"
    N, stat, mass, Ny, steps, meas, step = 32, 2, 2.0, 3, 12, 3, 0.0004
    out = f"out_{engine}_{precision}_{N}_{N}_mpi1_r0.dat"
    full, part = tmp_path / "full", tmp_path / "part"
    full.mkdir()
    part.mkdir()
    ck = "checkpoint_every = 4\n"
    write_cfg(full / "cfg", N, stat, mass, Ny, steps, meas, step, "CORRELATED_NOISE_COUPLING", engine,
              precision=precision, extra=ck)
    r = subprocess.run([DRIVER, "1", "1", "cfg"], cwd=full, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    # stopped inside sample 0 (after step 8, the checkpoint at 8), then inside sample 1
    for stop, resume in ((7, 0), (9, 1), (0, 1)):
        extra = ck + (f"stop_after = {stop}\n" if stop else "") + ("resume = 1\n" if resume else "")
        write_cfg(part / "cfg", N, stat, mass, Ny, steps, meas, step, "CORRELATED_NOISE_COUPLING", engine,
                  precision=precision, extra=extra)
        r = subprocess.run([DRIVER, "1", "1", "cfg"], cwd=part, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        assert os.path.exists(part / out) == (stop == 0)
    assert open(part / out).read() == open(full / out).read()


def test_driver_batch_list_uneven_sizes(jw, orc, tmp_path):
    """batch = 64:2,128:1: the size groups evolve in one batch; trajectory t (over the
    groups) has the noise key {seed, t}; each group writes its own file with its
    count as the stat normalisation."""
    assert os.path.exists(DRIVER), "driver not built"
    mass, Ny, steps, meas, step = 2.0, 3, 6, 2, 0.0004
    cfg = tmp_path / "cfg"
    write_cfg(cfg, 16, 1, mass, Ny, steps, meas, step, "FIXED_COUPLING", "fused",
              extra="batch = 64:2,128:1\nreunit_every = 3\n")
    r = subprocess.run([DRIVER, "1", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    for N, trajs in ((64, [0, 1]), (128, [2])):
        got = read_out(tmp_path / f"out_fused_fp64_{N}_{N}_mpi1_r0.dat")
        ref = oracle_output(orc, jw, N, len(trajs), mass, Ny, steps, meas, step, 3, trajs=trajs)
        assert got.shape == ref.shape and np.array_equal(got[:, :3], ref[:, :3])
        np.testing.assert_allclose(got[:, 3:], ref[:, 3:], rtol=2e-6, atol=1e-12)


@pytest.mark.parametrize("coupling,code", [("SQRT_COUPLING_CONSTANT", 0), ("NOISE_COUPLING_CONSTANT", 1)])
def test_driver_position_evolution(jw, orc, tmp_path, coupling, code):
    """EvolutionChoice = POSITION_EVOLUTION end to end (N = 16) against the oracle."""
    assert os.path.exists(DRIVER), "driver not built"
    N, stat, mass, Ny, steps, meas, step = 16, 1, 2.0, 3, 4, 2, 0.001
    cfg = tmp_path / "cfg"
    write_cfg(cfg, N, stat, mass, Ny, steps, meas, step, coupling, "reference")
    cfg.write_text(cfg.read_text().replace("MOMENTUM_EVOLUTION", "POSITION_EVOLUTION"))
    r = subprocess.run([DRIVER, "1", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = read_out(tmp_path / f"out_reference_fp64_{N}_{N}_mpi1_r0.dat")
    ref = oracle_output(orc, jw, N, stat, mass, Ny, steps, meas, step, code, position=True)
    assert got.shape == ref.shape and np.array_equal(got[:, :3], ref[:, :3])
    np.testing.assert_allclose(got[:, 3:], ref[:, 3:], rtol=2e-6, atol=1e-12)


def test_driver_rejects_multi_rank_grid(tmp_path):
    cfg = tmp_path / "cfg"
    write_cfg(cfg, 16, 1, 1.0, 1, 1, 1, 0.0004, "FIXED_COUPLING", "fused")
    r = subprocess.run([DRIVER, "2", "1", str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "1 x 1" in r.stderr
