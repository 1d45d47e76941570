"""The reference's RPI driver links against the drop-in modules (container side).

oracle/ref/build_dropin.sh compiles the reference's own rpi_ser.f90 and links it with the
drop-in `mcmod_mass` / `rp_bridge` / `instantonmod` / `verletint` of pimd-tunneling_amd/fortran
and libpimdtunnel.so: once with the device plugin, once with a host plugin written to the
reference's plugin API (mcmod_waterdimer.f90:1-105).  Every routine and module variable
rpi_ser.f90 uses must therefore exist with the reference's interface: get_align,
align_atoms, rotate_atoms, read_path, splint, instanton, UM, detJ, eucliddist, gauleg,
path / Vpath / splinepath / lampath, xtilde, well1 / well2, mass, betan, pi.  Skipped where
the reference's sources are absent (they are not part of this repository)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "rpi_ser.f90")),
                    reason="the reference's sources are not present")
@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "pimd-tunneling_amd", "lib", "libpimdtunnel.so")),
                    reason="libpimdtunnel.so is not built")
def test_reference_rpi_driver_links_against_the_dropin():
    res = subprocess.run(["bash", os.path.join(ROOT, "oracle", "ref", "build_dropin.sh")],
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    for kind in ("device", "host"):
        exe = os.path.join(ROOT, "oracle", "_ref", f"rpi_dropin_{kind}")
        assert os.access(exe, os.X_OK), exe
