"""Configuration errors of the output and postprocessing keys end the `lbm`
CLI with the reference's TERMM status (255) and message before any device
work, so they run without a GPU."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LBM = os.path.join(ROOT, "lbm_amd", "bin", "lbm")
REFCFG = os.path.join(ROOT, "tests", "golden", "reference_configs")


def run_bad(tmp_path, edit):
    cfg = json.load(open(os.path.join(REFCFG, "couette__couette.json")))
    cfg["solver"]["output_dir"] = str(tmp_path / "out")
    edit(cfg["solver"])
    p = tmp_path / "case.json"
    p.write_text(json.dumps(cfg))
    return subprocess.run([LBM, str(p)], cwd=tmp_path, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("edit,message", [
    (lambda s: s.update(solution_interval=0), "solution_interval must be >= 1"),
    (lambda s: s.update(cellFilter={"cellFilter": "nearWall"}), "Unknown output filter nearWall"),
    (lambda s: s.update(cellFilter={"cellFilter": "targetLvl", "outputLvl": 0}),
     "Outputting a lvl below the partition lvl is not possible!"),
    (lambda s: s.update(cellFilter={"level": 3}), "Invalid output configuration missing \"cellFilter\" key"),
    (lambda s: s.update(postprocessing={"p": {"type": "plane", "execute": "atEnd"}}), "Invalid functype"),
    (lambda s: s.update(postprocessing={"p": {"type": "line", "execute": "atMidnight", "A": [0, 0], "B": [1, 1]}}),
     "Invalid hook"),
    (lambda s: s.update(postprocessing={"p": {"type": "line", "execute": "atEnd", "A": [0, 0, 0], "B": [1, 1, 1]}}),
     "Invalid line end points"),
])
def test_bad_output_keys_fail_like_the_reference(edit, message, tmp_path):
    r = run_bad(tmp_path, edit)
    assert r.returncode == 255, r.stderr[-1500:]
    assert message in r.stderr, r.stderr[-1500:]
