"""Lattice-state layout (IonLayout), compensated FP32 value format
(pk_pack / pk_unpack) and Yee tile order (band_tile) of
php_fdtd_amd/csrc/phpfdtd_internal.h, checked on the
host by tests/layout_check.cpp, which includes the header the kernels and the
state I/O use:

  * for every atom count 1..32 and both state forms, each dof of each slot of
    two tiles has its own element inside its tile, a lane's group of dofs is
    one 16-byte vector and the 64 lanes' vectors one 1 KiB row;
  * compensated form: each dof has its own 4-bit compensation code, in the
    rows after the word rows of its tile, eight groups per 16-byte lane
    vector (two per dword, four codes per 16-bit half); the tile is 1 KiB
    per word row plus 1 KiB per eight of them;
  * when r's last code row holds 1-4 groups (and only then) v's codes of
    those groups sit in its dwords 2 and 3, next to r's codes of the same
    groups, on no other code, and v's tile is one row shorter; v's words and
    other codes sit where r's sit in r's tile;
  * r_decode inverts r_index and rejects padding dofs, compensation rows and
    slots past the end;
  * pk: zero is exact; the word is fl32(x); a value comes back within half a
    unit of 2^-4 ulp (2^-28 relative), one unit (2^-27) where a +-8
    correction is clamped to +-7; re-packing a stored value is the identity;
    a code read or written at a nibble index (or at a bit of a dword) is
    sign-extended and leaves its neighbour as it was;
  * band_tile (work-group -> tile of the single-domain Yee launches): over a
    launch of a multiple of 8 g work-groups every tile runs exactly once, and
    each group of g tiles on one XCD, the groups going to the XCDs in turn.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_layout_and_value_format(tmp_path):
    exe = str(tmp_path / "layout_check")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", os.path.join(HERE, "layout_check.cpp"), "-o",
                    exe], check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
