"""The process-wide launch-shape cache (host code of the C-ABI library, no
GPU): keys, admissibility checks, the text file and its earlier three-integer
format.  The reference fixes the corresponding choices per build with -D
macros (src/numerics_rtc.hpp:105-126); there is nothing of it to compare to,
so these tests pin the cache's own contract."""
import pytest


@pytest.fixture
def tune():
    from compact_schemes_amd import tune as t
    t.clear()
    yield t
    t.clear()


def test_line_entries_are_keyed_by_ends_and_precision(tune):
    tune.set_line(0, 256, 16, start=0, end=0, precision=0)
    tune.set_line(0, 256, 32, start=1, end=1, precision=0)
    tune.set_line(0, 256, 8, start=0, end=1, precision=2)
    assert tune.get_line(0, 256, 0, 0, 0) == 16
    assert tune.get_line(0, 256, 1, 1, 0) == 32
    assert tune.get_line(0, 256, 0, 1, 2) == 8
    assert tune.get_line(0, 256, 0, 1, 0) == 0          # other key: none
    assert tune.get_line(1, 256, 0, 0, 0) == 0          # the filter is its own operator
    tune.set_line(0, 256, 0, start=0, end=0)            # K = 0 clears one entry
    assert tune.get_line(0, 256, 0, 0, 0) == 0
    assert tune.get_line(0, 256, 1, 1, 0) == 32


def test_sweep_and_filter_entries(tune):
    assert tune.get_sweep(2, 256, 0, 0) is None
    tune.set_sweep(2, 256, 0, 0, 2, 16, 0)
    tune.set_sweep(2, 256, 0, 1, 0, 8, 1)
    assert tune.get_sweep(2, 256, 0, 0) == (2, 16, 0)
    assert tune.get_sweep(2, 256, 0, 1) == (0, 8, 1)
    assert tune.get_sweep(2, 256, 1, 0) is None
    assert tune.get_sweep(2, 256, 0, 0, metrics=1) is None     # per-point metrics: a key of its own
    tune.set_sweep(2, 256, 0, 0, 1, 8, 0, metrics=1)
    assert tune.get_sweep(2, 256, 0, 0, metrics=1) == (1, 8, 0)
    assert tune.get_sweep(2, 256, 0, 0) == (2, 16, 0)
    tune.set_filter_sweep(1, 128, 1, 5, 8, 16)
    assert tune.get_filter_sweep(1, 128, 1) == (5, 8, 16)
    assert tune.get_filter_sweep(1, 128, 0) is None
    tune.clear()
    assert tune.get_sweep(2, 256, 0, 0) is None and tune.get_filter_sweep(1, 128, 1) is None


@pytest.mark.parametrize("call", [
    lambda t: t.set_line(2, 64, 8),                 # no such operator
    lambda t: t.set_line(0, 64, 12),                # chunk length
    lambda t: t.set_line(0, 64, 8, start=2),        # end code
    lambda t: t.set_line(0, 64, 8, precision=5),
    lambda t: t.set_line(0, 4, 8),                  # shorter than the stencil
    lambda t: t.set_sweep(3, 64, 0, 0, 0, 16, 0),   # axis
    lambda t: t.set_sweep(1, 64, 0, 2, 0, 16, 0),   # precision
    lambda t: t.set_sweep(1, 64, 0, 0, 0, 16, 0, metrics=2),
    lambda t: t.set_sweep(1, 64, 0, 0, 3, 16, 0),   # variant
    lambda t: t.set_sweep(1, 64, 0, 0, 0, 12, 0),   # lines
    lambda t: t.set_sweep(1, 1024, 0, 0, 0, 16, 0), # longer than a fused axis
    lambda t: t.set_filter_sweep(1, 64, 0, 3, 16, 8),
    lambda t: t.set_filter_sweep(1, 64, 0, 5, 16, 32),
])
def test_inadmissible_entries_rejected(tune, call):
    import compact_schemes_amd as cs
    with pytest.raises(cs.CsError):
        call(tune)


def test_chunk_length_too_short_for_line_rejected(tune):
    """8-point chunks of a 1024-point line are 128 chunks, more than a line
    kernel takes."""
    import compact_schemes_amd as cs
    with pytest.raises(cs.CsError, match="too short"):
        tune.set_line(0, 1024, 8)
    tune.set_line(0, 1024, 16)
    assert tune.get_line(0, 1024) == 16


def test_save_load_round_trip(tune, tmp_path):
    entries = {
        "line": [(0, 96, 0, 0, 0, 16), (1, 96, 1, 1, 0, 32), (0, 256, 0, 1, 2, 8)],
        # axis, length, seam, precision, metrics, variant, lines, xcd
        "sweep": [(0, 256, 0, 0, 0, 1, 8, 0), (2, 256, 1, 1, 0, 2, 16, 1), (2, 256, 1, 1, 1, 0, 16, 0)],
        "filter": [(1, 256, 0, 1, 16, 16), (0, 512, 1, 5, 16, 8)],
    }
    for e in entries["line"]:
        tune.set_line(e[0], e[1], e[5], e[2], e[3], e[4])
    for e in entries["sweep"]:
        tune.set_sweep(*e[:4], *e[5:], metrics=e[4])
    for e in entries["filter"]:
        tune.set_filter_sweep(*e)
    path = tmp_path / "tune.txt"
    tune.save(path)
    text = path.read_text().splitlines()
    body = sorted(tuple(ln.split()) for ln in text if ln and not ln.startswith("#"))
    want = sorted((k,) + tuple(str(v) for v in e) for k, es in entries.items() for e in es)
    assert body == want
    tune.clear()
    tune.load(path)
    for e in entries["line"]:
        assert tune.get_line(e[0], e[1], e[2], e[3], e[4]) == e[5]
    for e in entries["sweep"]:
        assert tune.get_sweep(*e[:4], metrics=e[4]) == e[5:]
    for e in entries["filter"]:
        assert tune.get_filter_sweep(*e[:3]) == e[3:]
    tune.save(tmp_path / "again.txt")                   # load then save: the same file
    assert (tmp_path / "again.txt").read_text() == path.read_text()


def test_load_earlier_three_integer_format(tune, tmp_path):
    """A line "<op> <length> <K>" sets K for FP64 lines with closure rows at
    both ends and with seam rows at both ends."""
    path = tmp_path / "old.txt"
    path.write_text("# op length K\n0 96 16\n1 200 32\n\n")
    tune.load(path)
    for s in (0, 1):
        assert tune.get_line(0, 96, s, s, 0) == 16
        assert tune.get_line(1, 200, s, s, 0) == 32
    assert tune.get_line(0, 96, 0, 1, 0) == 0
    assert tune.get_line(0, 96, 0, 0, 1) == 0


def test_load_sweep_entry_without_metrics_field(tune, tmp_path):
    """A sweep line of seven integers (no metrics field) is an entry for
    constant metrics."""
    path = tmp_path / "sweep7.txt"
    path.write_text("sweep 1 128 0 1 2 8 1\n")
    tune.load(path)
    assert tune.get_sweep(1, 128, 0, 1) == (2, 8, 1)
    assert tune.get_sweep(1, 128, 0, 1, metrics=1) is None


@pytest.mark.parametrize("text, where", [
    ("line 0 64 0 0 0\n", ":1:"),                           # five integers
    ("sweep 0 64 0 0 1 8\n", ":1:"),                        # six, a sweep takes seven or eight
    ("sweep 0 64 0 0 0 1 8 0 1\n", ":1:"),                  # nine
    ("sweep 0 64 0 0 2 1 8 0\n", ":1:"),                    # metrics 2
    ("# ok\nline 0 64 0 0 0 8\nbogus 1 2\n", ":3:"),         # unknown entry
    ("filter 1 64 0 5 16 24\n", ":1:"),                     # chunk length 24
    ("line 0 64 0 0 0 8\n7 8\n", ":2:"),                    # two integers
])
def test_load_errors_name_file_and_line(tune, tmp_path, text, where):
    import compact_schemes_amd as cs
    path = tmp_path / "bad.txt"
    path.write_text(text)
    with pytest.raises(cs.CsError) as err:
        tune.load(path)
    assert str(path) + where in str(err.value)


def test_load_missing_file(tune, tmp_path):
    import compact_schemes_amd as cs
    with pytest.raises(cs.CsError, match="cannot read"):
        tune.load(tmp_path / "absent.txt")
