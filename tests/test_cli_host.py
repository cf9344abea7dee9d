"""Host-side behaviour of the drop-in tools, before any device work: usage and
exit code 1 on a wrong argument count (src/main/clique_PA.cu:35-47,
src/main/motifs_DM_WCLB.cu:39-51), and exit code 1 with a message for a
missing graph file (where the reference's unchecked fopen crashes,
src/main/Graph.cpp:25-29) or a motif size without a class table."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "dumato_amd", "bin")


def _run(tool, *args):
    return subprocess.run([os.path.join(BIN, tool), *map(str, args)], capture_output=True, text=True,
                          timeout=60)


@pytest.mark.parametrize("tool", ["clique_PA", "motifs", "q4_PA", "q13_PA"])
def test_wrong_argc_prints_usage(tool):
    r = _run(tool, "g.edgelist", 4)
    assert r.returncode == 1
    assert r.stdout.startswith("Usage: ")
    assert "Wrong amount of parameters!" in r.stdout


@pytest.mark.parametrize("tool", ["clique_PA", "motifs", "q5_PA"])
def test_missing_graph_file(tool, tmp_path):
    r = _run(tool, tmp_path / "absent.edgelist", 4, 102400, 256, 16, 30)
    assert r.returncode == 1
    assert "absent.edgelist" in r.stderr


def test_motif_size_without_table(tmp_path):
    g = tmp_path / "g.edgelist"
    g.write_text("0 1\n1 2\n")
    r = _run("motifs", g, 9, 102400, 256, 16, 30)
    assert r.returncode == 1
    assert "Generate file ../dictionaries/9.csv first" in r.stdout
