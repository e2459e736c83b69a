"""The committed profiler summaries are text, each with the command that made
it, and that command runs committed files only."""
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _summaries():
    return sorted(glob.glob(os.path.join(ROOT, "profiles", "*")))


def test_profiles_are_short_text_summaries():
    files = _summaries()
    assert files, "no profiler summaries"
    for f in files:
        assert f.endswith((".md", ".txt")), f
        assert os.path.getsize(f) < 16 * 1024, f


def test_each_summary_gives_its_rocprofv3_command():
    for f in _summaries():
        text = open(f).read()
        cmds = [l for l in text.splitlines() if "rocprofv3" in l and "--" in l]
        assert cmds, f"{os.path.basename(f)}: no rocprofv3 command"
        assert "_work/" not in text, f"{os.path.basename(f)} names a working file"
        # every script the commands run is in the repository
        for script in re.findall(r"\$R/([\w./-]+\.py)", text):
            assert os.path.exists(os.path.join(ROOT, script)), (f, script)
