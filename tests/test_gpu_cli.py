"""The command-line driver (`sevc -c config.ini`, in place of main.cu) on
the GPU: snapshots at the reference's schedule (main.cu:328), in the
export_data format, whose vertex coordinates equal the oracle's after the
same steps, with and without the in-step checks of debug_lvl = 1."""
import os
import subprocess

import numpy as np
import pytest

import sevc
from _util import REF2K, Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("debug_lvl", [0, 1])
def test_cli_snapshots_match_oracle(tmp_path, debug_lvl):
    cli = sevc.CLI_PATH
    assert os.path.exists(cli), "CLI not built"
    out = tmp_path / "run"
    ini = tmp_path / "c.ini"
    ini.write_text("\n".join([
        "[Temporal]", "dt=1e-4", "dtau=1e-4", "nmax=60", "snap_gap=20", "leave_percent=0.05",
        "[Grains]", "n_grains=2000",
        f"vertex_fname={REF2K}/vertices.txt", f"boundary_fname={REF2K}/ridges.txt",
        f"orientation_fname={REF2K}/oris.txt", f"SE_fname={REF2K}/SE/uniform.txt",
        "GB_eps=0.0", "GB_scaling=1.0", "SE_eps=0.5", "domain_bound=1.0", "do_nucleation=false",
        f"debug_lvl={debug_lvl}",
        "[Output]", f"folder={out}", ""]))
    print("  running the CLI for 61 steps", flush=True)
    r = subprocess.run([cli, "-c", str(ini)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    files = sorted(os.listdir(out / "DATA"))
    assert files == [f"{k:06d}.txt" for k in range(4)], files
    o = sevc.options_from_ini(str(ini))
    mesh = sevc.load_files(o)
    orc = Oracle(o, mesh)
    rc, st = orc.step(1000)
    assert rc == 0 and st.n == 61 and st.done
    ref = orc.get_mesh()
    lines = (out / "DATA" / "000003.txt").read_text().splitlines()
    start = lines.index(next(x for x in lines if x.startswith("#VERTEX"))) + 1
    xy = np.array([[float(t) for t in lines[start + i].split()[3:5]] for i in range(mesh.nv)])
    want = np.array([[float("%.16f" % v) for v in p] for p in ref.xy()])
    assert np.array_equal(xy, want)
    assert f"CURR GRAINS {st.n_grains}" in lines[:25]
