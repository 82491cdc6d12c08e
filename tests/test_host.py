"""Host-side checks (no GPU): the C ABI library loads and exports every symbol
include/jimwlk.h declares, argument errors, the host-only observables
post-processing, and the C++ host classes (config parser, rand_class, momenta)
built with the host compiler against the oracle."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "jimwlk.h")
HOSTDIR = os.path.join(REPO, "softx-d-21-00092_amd", "host")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jw_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(jw):
    L = jw.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_create_rejects_bad_sizes(jw):
    L = jw.lib()
    h = ctypes.c_void_p()
    for N in (0, 3, 100, 16384):
        assert L.jw_create(ctypes.byref(h), N, 0) == -1
        assert b"power of two" in L.jw_last_error()
    assert L.jw_version().startswith(b"jimwlk")


def test_radial_bins_and_qs_host(jw, orc):
    N = 32
    x = np.arange(N)
    d = np.minimum(x, N - x)
    d[N // 2] = N // 2
    r2 = (d[:, None] ** 2 + d[None, :] ** 2).astype(float)
    Qs_true = 0.35
    S = np.exp(-r2 * Qs_true ** 2 / 4.0).reshape(-1)      # S(R_s) = e^-1/2 at R_s = sqrt2 / Qs
    r, s, c = jw.radial_bins(S, N, N // 2)
    ro, so, co = orc.radial_bins(S, N, N // 2)
    assert np.array_equal(c, co) and np.allclose(r, ro, atol=1e-14) and np.allclose(s, so, atol=1e-14)
    q = jw.saturation_scale(r, s, c)
    assert abs(q - Qs_true) / Qs_true < 0.05
    # never below e^-1/2 -> 0
    assert jw.saturation_scale(r, np.ones_like(s), c) == 0.0


@pytest.fixture(scope="module")
def host_probe(tmp_path_factory):
    """Host-compiler build of a probe over the host classes (no GPU, no library)."""
    d = tmp_path_factory.mktemp("hostprobe")
    src = d / "probe.cpp"
    src.write_text(r'''
#include <stdio.h>
#include "config.h"
#include "lattice.h"
int main(int argc, char** argv) {
    config c;
    if (!c.read_config_from_file(argv[1])) return 1;
    printf("%d %.17g %.17g %d %s %d %d %.17g %d %d %d %d %u %d %d %d\n", c.stat, c.mu, c.mass,
           c.elementaryWilsonLines, c.file_name.c_str(), c.langevin_steps, c.measurements, c.step,
           (int)c.EvolutionChoice, (int)c.CouplingChoice, (int)c.KernelChoice, c.Nx, c.seed, (int)c.engine,
           (int)c.mixed, c.reunit_every);
    mpi_class mpi;
    mpi.mpi_init(&c);
    rand_class r(&mpi, &c);
    for (int i = 0; i < 6; i++) printf("%.17g ", r.get());
    printf("\n");
    momenta m(&c, &mpi);
    m.set();
    for (int i = 0; i < c.Nx * c.Ny; i += 37) printf("%.17g %.17g %.17g ", m.phat2(i), m.pbarX(i), m.pbarY(i));
    printf("\n");
    printf("%d %d %d %d %d", (int)c.float_parse, c.mv_zero_mode, c.checkpoint_every, (int)c.resume, c.stop_after);
    for (auto& g : c.batch) printf(" %d:%d", g.first, g.second);
    printf("\n");
    return 0;
}
''')
    exe = d / "probe"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{HOSTDIR}", str(src), "-o", str(exe)], check=True)
    return exe


def run_probe(exe, cfg, extra=False):
    out = subprocess.run([str(exe), str(cfg)], check=True, capture_output=True, text=True).stdout.split("\n")
    res = out[0].split(), [float(v) for v in out[1].split()], [float(v) for v in out[2].split()]
    return res + (out[3].split(),) if extra else res


def test_config_parser_run_keys(host_probe, tmp_path):
    """float_parse, mv_zero_mode, batch lists, checkpoint_every / resume / stop_after."""
    cfg = tmp_path / "cfg"
    base = open(os.path.join(REPO, "tests", "golden", "config_reference")).read()
    cfg.write_text(base)
    f, _, _, x = run_probe(host_probe, cfg, extra=True)
    assert x == ["1", "0", "0", "0", "0"]                   # defaults: the reference's float parse
    cfg.write_text(base + "float_parse = 0\nmv_zero_mode = 1\ncheckpoint_every = 25\nresume = 1\nstop_after = 60\n")
    f, _, _, x = run_probe(host_probe, cfg, extra=True)
    # mu, mass and step read as doubles: the file's decimal values to the last bit
    assert float(f[1]) == 30.72 and float(f[2]) == 0.0001 and float(f[7]) == 0.0004
    assert x == ["0", "1", "25", "1", "60"]
    cfg.write_text(base + "batch = 256:48,512:32,1024:16\n")
    f, _, _, x = run_probe(host_probe, cfg, extra=True)
    assert x[5:] == ["256:48", "512:32", "1024:16"] and int(f[13]) == 2     # the batched engine
    for bad in ("batch = 256\n", "batch = 256:0\n", "batch = 64:2;128:1\n", "mv_zero_mode = 2\n",
                "checkpoint_every = -1\n"):
        cfg.write_text(base + bad)
        assert subprocess.run([str(host_probe), str(cfg)], capture_output=True).returncode == 1, bad


def test_config_parser_reference_file(host_probe):
    f, _, _ = run_probe(host_probe, os.path.join(REPO, "tests", "golden", "config_reference"))
    f32 = lambda v: float(np.float32(v))
    assert int(f[0]) == 2
    # mu, mass and step pass through a float (config.h:79-93)
    assert float(f[1]) == f32(30.72) and float(f[2]) == f32(0.0001) and float(f[7]) == f32(0.0004)
    assert float(f[7]) != 0.0004
    assert int(f[3]) == 50 and f[4] == "output" and int(f[5]) == 100 and int(f[6]) == 100
    assert (int(f[8]), int(f[9]), int(f[10])) == (1, 1, 1)   # MOMENTUM, NOISE, SIN
    assert int(f[11]) == 64 and int(f[13]) == 0               # defaults: N = 64, fused engine
    assert int(f[14]) == 0 and int(f[15]) == 10               # FP64; projection every 10 (mixed only)


def test_config_parser_extra_keys(host_probe, tmp_path):
    cfg = tmp_path / "cfg"
    base = open(os.path.join(REPO, "tests", "golden", "config_reference")).read()
    base = base.replace("NOISE_COUPLING_CONSTANT", "CORRELATED_NOISE_COUPLING")
    cfg.write_text(base + "N = 16\nseed = 42\nengine = reference\n")
    f, _, mom = run_probe(host_probe, cfg)
    assert int(f[9]) == 4 and int(f[11]) == 16 and int(f[12]) == 42 and int(f[13]) == 1
    cfg.write_text(base + "engine = batched\nprecision = mixed\nreunit_every = 5\n")
    f, _, _ = run_probe(host_probe, cfg)
    assert int(f[13]) == 2 and int(f[14]) == 1 and int(f[15]) == 5
    for bad in ("engine = turbo\n", "precision = fp16\n", "engine = reference\nprecision = mixed\n"):
        cfg.write_text(base + bad)
        assert subprocess.run([str(host_probe), str(cfg)], capture_output=True).returncode == 1, bad


def test_host_rand_class_and_momenta(host_probe, orc, tmp_path):
    cfg = tmp_path / "cfg"
    cfg.write_text(open(os.path.join(REPO, "tests", "golden", "config_reference")).read() + "N = 16\nseed = 7\n")
    _, rnd, mom = run_probe(host_probe, cfg)
    # rand_class: Philox4x32-10, key {seed, 0xFFFFFFFF}, counter {c lo, c hi, 0, 0}, two words per uniform
    expect = []
    for c in range(3):
        o = orc.philox4x32_10([c, 0, 0, 0], [7, 0xFFFFFFFF])
        for a, b in ((o[0], o[1]), (o[2], o[3])):
            expect.append((float(((a << 32) | b) >> 11) + 0.5) * 2.0 ** -53)
    assert rnd == expect
    assert all(0.0 < v < 1.0 for v in rnd)
    ph, bx, by = orc.momenta(16)
    idx = list(range(0, 256, 37))
    got = np.array(mom).reshape(-1, 3)
    assert np.array_equal(got[:, 0], ph[idx]) and np.array_equal(got[:, 1], bx[idx]) and np.array_equal(got[:, 2], by[idx])
