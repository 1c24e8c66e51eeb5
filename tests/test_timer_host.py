"""spla_timer_* (SPLA src/timing/timing.cpp:49-86) through the C ABI, on the host.

The timer needs no GPU: nesting per thread, accumulation of repeated scopes,
the JSON layout {"name": {"timings": [...], "sub-timings": {...}}}, the printed
table, unmatched stops and argument errors.
"""

from __future__ import annotations

import json
import threading
import time

import pytest

import spla_amd as spla
from spla_amd import _lib


@pytest.fixture()
def timer(tmp_path):
    spla.timer.clear()
    yield spla.timer
    spla.timer.clear()
    spla.timer.set_library_scopes(False)


def exported(timer, tmp_path):
    path = tmp_path / "t.json"
    timer.export_json(str(path))
    return json.loads(path.read_text())


def test_nested_scopes_accumulate(timer, tmp_path):
    timer.start("outer")
    for _ in range(3):
        timer.start("inner")
        time.sleep(0.01)
        timer.stop("inner")
    timer.start("other")
    timer.stop("other")
    timer.stop("outer")
    timer.start("outer")
    timer.stop("outer")
    t = exported(timer, tmp_path)
    assert list(t) == ["outer"]
    outer = t["outer"]
    assert len(outer["timings"]) == 2
    sub = outer["sub-timings"]
    assert list(sub) == ["inner", "other"]  # order of first appearance
    inner = sub["inner"]["timings"]
    assert len(inner) == 3 and all(x >= 0.01 for x in inner)
    assert sub["inner"]["sub-timings"] == {}
    # the children of the first outer scope lie inside it
    assert sum(inner) + sub["other"]["timings"][0] <= outer["timings"][0]


def test_table_lists_every_scope(timer, capfd):
    timer.start("solve")
    timer.start("step")
    timer.stop("step")
    timer.start("step")
    timer.stop("step")
    timer.stop("solve")
    capfd.readouterr()
    timer.print()
    out = capfd.readouterr().out
    header = [line for line in out.splitlines() if "Parent %" in line]
    assert len(header) == 1 and "Median" in header[0] and "Max" in header[0]
    rows = {line.split()[0]: line for line in out.splitlines() if line.strip().startswith(("solve", "step"))}
    assert rows["solve"].split()[1] == "1"
    assert rows["step"].split()[1] == "2"
    assert any(line.startswith("  step") for line in out.splitlines())  # indented child


def test_unmatched_stop_is_a_warning(timer, tmp_path, capfd):
    timer.stop("never-started")  # SPLA's timer accepts it and reports it when processing
    timer.start("a")
    timer.stop("b")
    timer.stop("a")
    t = exported(timer, tmp_path)
    assert list(t) == ["a"] and len(t["a"]["timings"]) == 1
    capfd.readouterr()
    timer.print()
    out = capfd.readouterr().out
    assert out.count("does not match the innermost open scope") == 2


def test_threads_have_their_own_nesting(timer, tmp_path):
    barrier = threading.Barrier(2)

    def work(name):
        timer.start(name)
        barrier.wait()  # both scopes open at once
        timer.start("leaf")
        timer.stop("leaf")
        barrier.wait()
        timer.stop(name)

    threads = [threading.Thread(target=work, args=(n,)) for n in ("t0", "t1")]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    t = exported(timer, tmp_path)
    assert sorted(t) == ["t0", "t1"]
    for n in ("t0", "t1"):
        assert list(t[n]["sub-timings"]) == ["leaf"]


def test_names_are_escaped_in_json(timer, tmp_path):
    name = 'quote " back \\ tab \t'
    timer.start(name)
    timer.stop(name)
    t = exported(timer, tmp_path)
    assert list(t) == [name]


def test_clear_drops_open_scopes(timer, tmp_path):
    timer.start("open")
    timer.clear()
    timer.stop("open")  # belongs to the dropped tree: a warning only
    timer.start("fresh")
    timer.stop("fresh")
    assert list(exported(timer, tmp_path)) == ["fresh"]


def test_argument_errors(timer, tmp_path):
    lib = _lib.lib()
    assert lib.spla_timer_start(-1, b"x") == spla.SplaError.INVALID_PARAMETER_ERROR
    assert lib.spla_timer_stop(3, None) == spla.SplaError.INVALID_POINTER_ERROR
    # (length, pointer): only the first n bytes name the scope
    assert lib.spla_timer_start(3, b"abcdef") == 0
    assert lib.spla_timer_stop(3, b"abcXYZ") == 0
    assert list(exported(timer, tmp_path)) == ["abc"]
    with pytest.raises(spla.InvalidParameterError):
        timer.export_json(str(tmp_path / "missing-dir" / "t.json"))


def test_library_scopes_are_off_by_default(timer, tmp_path):
    # no product runs here; the switch itself is checked on the GPU
    timer.set_library_scopes(True)
    timer.set_library_scopes(False)
    assert exported(timer, tmp_path) == {}
