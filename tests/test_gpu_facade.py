"""The C++ drop-in (include/framegen.h) driven by a host program (tests/cpp/facade_check, built with the library):
the frames it writes equal the oracle's, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from fgr import scenes
from oracle import v1

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "cpp", "facade_check")


def test_cpp_facade_matches_the_oracle(tmp_path):
    assert os.path.exists(EXE), "tests/cpp/facade_check is built by __graft_entry__.build() (make)"
    a, c = scenes.pair_translate(51, w=320, h=180, sprite=40, sprite_pos=(120, 60), shift=(4, -2))
    pa, pc, po = tmp_path / "a.rgba", tmp_path / "c.rgba", str(tmp_path / "out")
    a.tofile(pa)
    c.tofile(pc)
    ts = ("0.5", "0.25", "0.75")
    r = subprocess.run([EXE, "320", "180", str(pa), str(pc), po, *ts], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    est = v1.estimate(a, c)
    assert f"sad blocks {est['nblocks']}" in r.stdout
    for i, t in enumerate(ts):
        got = np.fromfile(f"{po}_{i}.rgba", np.uint8).reshape(180, 320, 4)
        want, _, _ = v1.synth(a, c, est, float(t))
        assert np.array_equal(got, want), t
