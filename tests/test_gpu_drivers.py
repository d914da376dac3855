"""The C++ drop-in block (SWE_DimensionalSplittingHip) driven by the example
drivers, against the oracle: single block (swe_simple's loop), many blocks
with global time stepping, and many blocks with local time stepping (the
overdecomposed driver's loop and the LTS state machine)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import same_bits

import snapshots

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "swe-ppm_amd" / "host" / "build"


def run(args, timeout=120):
    print("running", " ".join(str(a) for a in args), flush=True)
    p = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr + p.stdout
    return p.stdout


@pytest.mark.parametrize("solver,flavour", [("augrie", "u"), ("fwave", "u"), ("hybrid", "s")])
def test_swe_simple_driver_matches_oracle(oracle, tmp_path, solver, flavour):
    import swe_ppm_amd as swe

    i0, f1 = tmp_path / "init.bin", tmp_path / "final.bin"
    out = run([BIN / "swe_simple_hip", "-t", "100", "-n", "1", "-x", "96", "-y", "80", "--solver", solver,
               "--flavour", flavour, "--scenario", "radial", "--max-steps", "40", "--dump-initial", i0, "--dump", f1,
               "-w", "1", "-o", tmp_path / "vtk"])
    assert "iterations 40" in out
    s0, s1 = snapshots.read(i0)[0], snapshots.read(f1)[0]
    dx, dy = 1000.0 / 96, 1000.0 / 80
    bnd = [0, 0, 0, 0]
    sid = swe.SOLVER_NAMES[solver]
    if flavour == "u":
        oh, ohu, ohv, _ = oracle.run_u(s0["h"], s0["hu"], s0["hv"], s0["b"], np.float32(dx), np.float32(dy), bnd,
                                       sid, 40)
    else:
        oh, ohu, ohv, _ = oracle.run_s(s0["h"], s0["hu"], s0["hv"], s0["b"], np.float32(dx), bnd, sid, 40)
    for a, c in ((s1["h"], oh), (s1["hu"], ohu), (s1["hv"], ohv)):
        assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1])
    # the VTK writer produced the t=0 and the checkpoint files
    vts = sorted(tmp_path.glob("vtk.*.vts"))
    assert len(vts) == 2 and "<DataArray Name=\"h\"" in vts[1].read_text()


def _oracle_blocks(states, rows, bnd_default=0):
    n = len(states)
    cols = n // rows
    out = []
    for k, s in enumerate(states):
        c, r = divmod(k, rows)
        nbr = [k - rows if c > 0 else -1, k + rows if c < cols - 1 else -1, k - 1 if r > 0 else -1,
               k + 1 if r < rows - 1 else -1]
        bnd = [4 if m >= 0 else bnd_default for m in nbr]
        out.append(dict(nx=s["nx"], ny=s["ny"], bnd=bnd, nbr=nbr, h=s["h"], hu=s["hu"], hv=s["hv"], b=s["b"]))
    return out


def test_overdecomp_global_matches_oracle(oracle, tmp_path):
    """6 blocks of the reference's radial-bathymetry dam break, global dt."""
    i0, f1 = tmp_path / "init.bin", tmp_path / "final.bin"
    run([BIN / "swe_overdecomp_hip", "-t", "1000", "-n", "1", "-x", "150", "-y", "100", "-b", "6", "--solver",
         "augrie", "--max-steps", "30", "--dump-initial", i0, "--dump", f1])
    s0, s1 = snapshots.read(i0), snapshots.read(f1)
    rows = 2  # floor(sqrt(6)) = 2 divides 6
    blocks = _oracle_blocks(s0, rows)
    res, upd, its, dts = oracle.multiblock_run(blocks, np.float32(2000 / 150), np.float32(2000 / 100), 2, False,
                                              1.0, 1e30, 30)
    for k, (g, o) in enumerate(zip(s1, res)):
        for a, c in zip((g["h"], g["hu"], g["hv"]), o[:3]):
            assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), f"block {k}"


def test_overdecomp_lts_matches_oracle(oracle, tmp_path):
    """28 blocks of uneven size and depth (7 rows x 4 columns), local time
    stepping with macro step 1 s: the C++ drop-in blocks driven through the
    reference's LTS loop reproduce the oracle's LTS run."""
    rng = np.random.default_rng(3)
    widths, heights = [8, 16, 24, 12], [8, 12, 16, 8, 20, 12, 8]
    small = []
    for c in range(4):
        for r in range(7):
            nx, ny = widths[c], heights[r]
            h = np.full((nx + 2, ny + 2), np.float32(rng.uniform(2.0, 40.0)), np.float32)
            z = np.zeros_like(h)
            small.append(dict(nx=nx, ny=ny, h=h, hu=z.copy(), hv=z.copy(), b=z.copy()))
    i0, f1 = tmp_path / "init.bin", tmp_path / "final.bin"
    snapshots.write(i0, small)
    out = run([BIN / "swe_overdecomp_hip", "-t", "3", "-n", "1", "-l", "1", "--rows", "7", "--solver", "augrie",
               "--input", i0, "--dump", f1])
    s1 = snapshots.read(f1)
    ob = _oracle_blocks(small, 7)
    ntot_x = sum(small[c * 7]["nx"] for c in range(4))
    ntot_y = sum(small[r]["ny"] for r in range(7))
    res, upd, its, _ = oracle.multiblock_run(ob, np.float32(2000 / ntot_x), np.float32(2000 / ntot_y), 2, True,
                                            1.0, 3.0, 3)
    line = [ln for ln in out.splitlines() if ln.startswith("updates ")][0]
    gupd = [int(v) for v in line.split()[1:]]
    assert gupd == [int(u) for u in upd], (gupd, upd)
    for k, (g, o) in enumerate(zip(s1, res)):
        for a, c in zip((g["h"], g["hu"], g["hv"]), o[:3]):
            assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1]), f"block {k}"


def _assert_nc_matches(path, state, nrec, t_last):
    """The last record of a driver's netCDF file holds the interior of the
    state the driver dumped at the same instant, as (y, x) slabs."""
    from scipy.io import netcdf_file

    with netcdf_file(str(path), "r", mmap=False) as f:
        v = f.variables
        assert v["ah"].shape[0] == nrec
        assert np.float32(v["time"].data[0]) == 0.0
        assert abs(float(v["time"].data[-1]) - t_last) <= 1e-6 * max(1.0, t_last)  # printed with %f
        for name, key in (("ah", "h"), ("hu", "hu"), ("hv", "hv")):
            assert same_bits(v[name].data[-1], state[key][1:-1, 1:-1].T), name
        assert same_bits(v["b"].data, state["b"][1:-1, 1:-1].T)


def test_swe_simple_driver_writes_netcdf(tmp_path):
    f1 = tmp_path / "final.bin"
    out = run([BIN / "swe_simple_hip", "-t", "1000", "-n", "1", "-x", "48", "-y", "40", "--solver", "augrie",
               "--scenario", "radial", "--max-steps", "25", "--dump", f1, "-w", "1", "--writer", "netcdf", "-o",
               tmp_path / "run"])
    t = float([ln for ln in out.splitlines() if ln.startswith("Write timestep")][-1].split("(")[1].rstrip("s)"))
    _assert_nc_matches(tmp_path / "run.nc", snapshots.read(f1)[0], 2, t)


@pytest.mark.parametrize("writer", ["netcdf", "vtk"])
def test_async_snapshot_writer_gives_the_same_files(tmp_path, writer):
    """--async-write (device snapshots, double-buffered, written by a thread
    of their own while the loop runs on) writes the same bytes as the
    synchronous writer: 6 checkpoints, 2 snapshots in flight at most."""
    outs = {}
    for mode in ("sync", "async"):
        d = tmp_path / mode
        d.mkdir()
        args = [BIN / "swe_simple_hip", "-t", "60", "-n", "6", "-x", "130", "-y", "90", "--solver", "augrie",
                "--scenario", "radial", "-w", "1", "--writer", writer, "-o", d / "run"]
        run(args + (["--async-write"] if mode == "async" else []))
        outs[mode] = {p.name: p.read_bytes() for p in sorted(d.iterdir())}
    assert len(outs["sync"]) == (1 if writer == "netcdf" else 7)
    assert outs["sync"].keys() == outs["async"].keys()
    for name in outs["sync"]:
        assert outs["sync"][name] == outs["async"][name], name


def test_overdecomp_driver_writes_one_netcdf_file_per_block(tmp_path):
    f1 = tmp_path / "final.bin"
    out = run([BIN / "swe_overdecomp_hip", "-t", "1000", "-n", "1", "-x", "60", "-y", "40", "-b", "6", "--solver",
               "augrie", "--max-steps", "12", "--dump", f1, "-w", "1", "-o", tmp_path / "od"])
    t = float([ln for ln in out.splitlines() if ln.startswith("Write timestep")][-1].split("(")[1].rstrip("s)"))
    s1 = snapshots.read(f1)
    rows = 2
    for k, s in enumerate(s1):
        c, r = divmod(k, rows)
        _assert_nc_matches(tmp_path / f"od_{c}{r}.nc", s, 2, t)


def _tsunami_files(tmp_path):
    """A coast (bathymetry -800 m .. +40 m) on a 40 x 20 grid of 100 m cells
    and a 3 m sea-floor uplift on a 12 x 8 grid of 50 m cells."""
    from scipy.io import netcdf_file

    def write(path, xs, ys, z):
        with netcdf_file(str(path), "w", version=2) as f:
            f.createDimension("y", len(ys))
            f.createDimension("x", len(xs))
            f.createVariable("x", "f4", ("x",))[:] = xs
            f.createVariable("y", "f4", ("y",))[:] = ys
            f.createVariable("z", "f4", ("y", "x"))[:] = z

    xs, ys = 50.0 + 100.0 * np.arange(40), 50.0 + 100.0 * np.arange(20)
    X, Y = np.meshgrid(xs, ys)
    z = np.float32(-800.0) + np.float32(840.0) * (X / 4000.0) ** 3 + 20.0 * np.sin(Y / 300.0)
    dxs, dys = 1025.0 + 50.0 * np.arange(12), 725.0 + 50.0 * np.arange(8)
    DX, DY = np.meshgrid(dxs, dys)
    dz = 3.0 * np.exp(-((DX - 1300.0) ** 2 + (DY - 900.0) ** 2) / (2 * 150.0 ** 2))
    write(tmp_path / "bathy.nc", xs, ys, z.astype(np.float32))
    write(tmp_path / "displ.nc", dxs, dys, dz.astype(np.float32))
    return tmp_path / "bathy.nc", tmp_path / "displ.nc"


def test_file_scenario_sampled_on_device_equals_host_loop(oracle, tmp_path):
    """swe_hip_init_grid_scenario against SWE_Block::initScenario over the
    same SWE_NetCdfScenario (ghosts and boundary conditions included), then
    20 steps against the oracle."""
    bathy, displ = _tsunami_files(tmp_path)
    common = ["-t", "1000", "-n", "1", "-x", "60", "-y", "40", "--solver", "augrie", "-b", bathy, "-d", displ]
    dev0, dev1, host0 = tmp_path / "dev0.bin", tmp_path / "dev1.bin", tmp_path / "host0.bin"
    run([BIN / "swe_simple_hip"] + common + ["--max-steps", "20", "--dump-initial", dev0, "--dump", dev1])
    run([BIN / "swe_simple_hip"] + common + ["--max-steps", "1", "--host-init", "--dump-initial", host0])
    d0, h0 = snapshots.read(dev0)[0], snapshots.read(host0)[0]
    for k in ("h", "hu", "hv", "b"):
        assert same_bits(d0[k].view(np.uint32), h0[k].view(np.uint32)), k
    assert (d0["h"] > 0).any() and (d0["h"] == 0).any()  # sea and dry land
    assert not same_bits(d0["b"] + d0["h"], np.zeros_like(d0["b"]))
    d1 = snapshots.read(dev1)[0]
    oh, ohu, ohv, _ = oracle.run_u(d0["h"], d0["hu"], d0["hv"], d0["b"], np.float32(4000 / 60), np.float32(2000 / 40),
                                   [0, 0, 0, 0], 2, 20)
    for a, c in ((d1["h"], oh), (d1["hu"], ohu), (d1["hv"], ohv)):
        assert same_bits(a[1:-1, 1:-1], c[1:-1, 1:-1])


ANALYTIC = ["radial", "centred", "radial-bathymetry", "bathymetry", "half-dry", "splashing-pool", "splashing-cone",
            "sea-at-rest"]


def test_analytic_scenarios_sampled_on_device_equal_host_loop(tmp_path):
    """swe_hip_init_scenario against SWE_Block::initScenario over the same
    scenario objects (ghosts and boundary conditions included), bit for bit:
    every analytic scenario of the single-block driver, and the six blocks at
    their own origins of the overdecomposed driver."""
    for name in ANALYTIC:
        dev, host = tmp_path / f"{name}_dev.bin", tmp_path / f"{name}_host.bin"
        common = [BIN / "swe_simple_hip", "-t", "1", "-n", "1", "-x", "70", "-y", "45", "--solver", "augrie",
                  "--scenario", name, "--max-steps", "1"]
        run(common + ["--dump-initial", dev])
        run(common + ["--host-init", "--dump-initial", host])
        d, h = snapshots.read(dev)[0], snapshots.read(host)[0]
        for k in ("h", "hu", "hv", "b"):
            assert same_bits(d[k].view(np.uint32), h[k].view(np.uint32)), (name, k)
    dev, host = tmp_path / "od_dev.bin", tmp_path / "od_host.bin"
    common = [BIN / "swe_overdecomp_hip", "-t", "1000", "-n", "1", "-x", "150", "-y", "100", "-b", "6", "--solver",
              "augrie", "--max-steps", "1"]
    run(common + ["--dump-initial", dev])
    run(common + ["--host-init", "--dump-initial", host])
    for k, (d, h) in enumerate(zip(snapshots.read(dev), snapshots.read(host))):
        for f in ("h", "hu", "hv", "b"):
            assert same_bits(d[f].view(np.uint32), h[f].view(np.uint32)), ("block", k, f)


def test_seeded_workloads_sampled_on_device_equal_host_generator():
    """swe_hip_init_seeded against swe_ppm_amd.scenarios' host sampling of
    the same parameter sets, bit for bit (ghosts and boundary conditions
    included): the 256^2 dam break, the 2048^2 islands, the 4096^2 trench,
    ensemble instances (dx != dy) and LTS blocks at their own origins with
    connected sides."""
    import swe_ppm_amd as swe
    from swe_ppm_amd import scenarios

    def check(name, nx, ny, dx, dy, bnd, spec, host, ox=0.0, oy=0.0):
        print("seeded workload", name, nx, ny, flush=True)
        with swe.SweBlock(nx, ny, dx, dy, ox, oy, solver=swe.AUGRIE) as blk:
            blk.set_boundaries(bnd)
            blk.init_seeded(spec)
            dev = blk.download(with_b=True)
        for f, d, e in zip(("h", "hu", "hv", "b"), dev, host):
            assert same_bits(d.view(np.uint32), np.asarray(e, np.float32).view(np.uint32)), (name, f)

    st = scenarios.dam_break_c1(256)
    check("c1", 256, 256, st[4], st[5], st[6], scenarios.dam_break_c1_spec(), st[:4])
    st = scenarios.dam_break_islands(n=2048, seed=2048)
    check("islands", 2048, 2048, st[4], st[5], st[6], scenarios.dam_break_islands_spec(2048), st[:4])
    st = scenarios.trench_hump(n=4096, seed=4096)
    check("trench", 4096, 4096, st[4], st[5], st[6], scenarios.trench_hump_spec(4096), st[:4])
    for k, i in enumerate(scenarios.ensemble_instances(count=6, seed=512)):
        check(f"ensemble {k}", i["nx"], i["ny"], i["dx"], i["dy"], i["bnd"], i["spec"],
              (i["h"], i["hu"], i["hv"], i["b"]))
    blocks, dx, dy = scenarios.lts_tiling(seed=28, n=512)
    for k, bl in enumerate(blocks[:6]):
        check(f"lts block {k}", bl["nx"], bl["ny"], dx, dy, bl["bnd"], bl["spec"],
              (bl["h"], bl["hu"], bl["hv"], bl["b"]), bl["ox"], bl["oy"])
    with swe.SweBlock(8, 8, 1.0, 1.0) as blk:
        with pytest.raises(swe.SweHipError):
            blk.init_seeded(dict(kind=7))


def _summary(out):
    upd = [ln for ln in out.splitlines() if ln.startswith("updates ")][0].split()[1:]
    head = [ln for ln in out.splitlines() if ln.startswith("blocks ")][0].split()
    return [int(u) for u in upd], int(head[head.index("iterations") + 1]), head[head.index("simulated") + 1]


def _assert_same_dumps(p, q):
    """Interiors bit for bit; b in full.  The ghost cells differ by design:
    the batched engine leaves the OUTFLOW / WALL ghosts and corners of the
    updated state, the block loop those of the state before the last
    update (the next setGhostLayer rewrites them in both)."""
    for k, (g, o) in enumerate(zip(snapshots.read(p), snapshots.read(q))):
        for name in ("h", "hu", "hv"):
            assert same_bits(g[name][1:-1, 1:-1].view(np.uint32), o[name][1:-1, 1:-1].view(np.uint32)), \
                f"block {k} {name}"
        assert same_bits(g["b"].view(np.uint32), o["b"].view(np.uint32)), f"block {k} b"


@pytest.mark.parametrize("lts", [False, True])
def test_overdecomp_batched_equals_block_loop(tmp_path, lts):
    """--batched (the whole loop in swe_hip_batch) against the block-by-block
    loop of the same driver: same states, update counts, iterations and
    checkpoint times."""
    if lts:
        rng = np.random.default_rng(11)
        small = []
        for c, nx in enumerate([16, 24, 8]):
            for r, ny in enumerate([12, 8, 20, 16]):
                h = np.full((nx + 2, ny + 2), np.float32(rng.uniform(2.0, 40.0)), np.float32)
                z = np.zeros_like(h)
                small.append(dict(nx=nx, ny=ny, h=h, hu=z.copy(), hv=z.copy(), b=z.copy()))
        i0 = tmp_path / "init.bin"
        snapshots.write(i0, small)
        args = ["-t", "4", "-n", "2", "-l", "1", "--rows", "4", "--solver", "augrie", "--input", i0]
    else:
        args = ["-t", "1000", "-n", "3", "-x", "90", "-y", "60", "-b", "6", "--solver", "hlle", "--max-steps", "25"]
    fa, fb = tmp_path / "loop.bin", tmp_path / "batched.bin"
    out_a = run([BIN / "swe_overdecomp_hip"] + args + ["--dump", fa])
    out_b = run([BIN / "swe_overdecomp_hip"] + args + ["--dump", fb, "--batched"])
    assert _summary(out_a) == _summary(out_b)
    assert [ln for ln in out_a.splitlines() if ln.startswith("Write")] == \
           [ln for ln in out_b.splitlines() if ln.startswith("Write")]
    _assert_same_dumps(fa, fb)
    if lts:
        assert len(set(_summary(out_a)[0])) > 1  # blocks took different numbers of sub-steps
