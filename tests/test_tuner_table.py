"""The tuner's table on the host: entries set, saved, cleared and loaded
back (no device calls)."""
import ctypes as C

from gcge_amd import abi


def read_all(lib):
    out = []
    i = 0
    while True:
        v = [C.c_int() for _ in range(5)]
        d = [C.c_double(), C.c_double()]
        if not lib.gcge_tune_entry(i, *[C.byref(x) for x in v], *[C.byref(x) for x in d]):
            return out
        out.append(tuple(x.value for x in v))
        i += 1


def test_table_round_trip(tmp_path):
    lib = abi.load()
    keep = str(tmp_path / "kept.txt")
    lib.gcge_tune_save(keep.encode())
    try:
        lib.gcge_tune_clear()
        assert read_all(lib) == []
        assert lib.gcge_tune_set(1, 8, 21, 5, 3) == 0
        assert lib.gcge_tune_set(2, 14, 1, 0, 7) == 0
        assert lib.gcge_tune_set(2, 14, 1, 0, 2) == 0      # replaces
        assert lib.gcge_tune_set(1, 8, 21, 5, -1) == -1    # rejected
        want = [(1, 8, 21, 5, 3), (2, 14, 1, 0, 2)]
        assert read_all(lib) == want
        path = tmp_path / "t.txt"
        assert lib.gcge_tune_save(str(path).encode()) == 2
        text = path.read_text()
        assert text.startswith("#")
        # comments, blank and malformed lines are skipped on load
        path.write_text(text + "\n# note\nnot a line\n1 16 10 0 1\n")
        lib.gcge_tune_clear()
        assert lib.gcge_tune_load(str(path).encode()) == 3
        assert read_all(lib) == [(1, 8, 21, 5, 3), (1, 16, 10, 0, 1), (2, 14, 1, 0, 2)]
        assert lib.gcge_tune_load(str(tmp_path / "missing.txt").encode()) == -1
        lib.gcge_tune_set_mode(1)
        assert lib.gcge_tune_get_mode() == 1
        lib.gcge_tune_set_mode(0)
        assert lib.gcge_tune_get_mode() == 0
    finally:
        lib.gcge_tune_clear()
        lib.gcge_tune_load(keep.encode())
