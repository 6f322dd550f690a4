"""Host-side parts of the Python mirror that need no GPU: the generator text parser of
read_generators / import_generators (src/tessellation.rs:126-171)."""
import numpy as np

from conftest import GOLDEN
from vorothree_amd import parse_generators


def test_parse_rules():
    # blank and whitespace-only lines are skipped (:157); the first token is the id (:162); x, y, z
    # are the next three tokens, 0.0 where missing (:164-167); later tokens are ignored
    text = "1 0.5 0.25 2\n\n \t \n7 1e-3\n8\n9 -4 5 6 extra tokens\n"
    assert np.array_equal(parse_generators(text), [[0.5, 0.25, 2.0], [1e-3, 0.0, 0.0], [0.0, 0.0, 0.0],
                                                   [-4.0, 5.0, 6.0]])
    assert parse_generators("").shape == (0, 3)


def test_parse_numbers_as_rust_does():
    # str::parse::<f64> takes signs, a leading or trailing point, exponents and inf/infinity/nan in
    # any case; it rejects underscores, hexadecimal and a lone point, which become 0.0 (:165)
    p = parse_generators("a +.5 5. -1E2\nb INF -Infinity nan\nc 1_0 0x10 .\nd 1e 2e+3 -0\n")
    assert np.array_equal(p[0], [0.5, 5.0, -100.0])
    assert p[1][0] == np.inf and p[1][1] == -np.inf and np.isnan(p[1][2])
    assert np.array_equal(p[2], [0.0, 0.0, 0.0])
    assert p[3][0] == 0.0 and p[3][1] == 2000.0 and p[3][2] == 0.0 and np.signbit(p[3][2])


def test_parse_lines_and_whitespace_as_rust_does():
    # str::lines splits at \n and drops one trailing \r; split_whitespace splits at Unicode
    # White_Space (tab, vertical tab, form feed, no-break space, ...) but not at \x1c-\x1f
    p = parse_generators("1\t2\x0b3\x0c4\r\n2\xa01 2 3\n3 1\x1c2 3 4\r")
    assert np.array_equal(p, [[2.0, 3.0, 4.0], [1.0, 2.0, 3.0], [0.0, 3.0, 4.0]])


def test_parse_the_voropp_input_files():
    # the reference's comparison tests read these files with import_generators (tests/comparisons.rs:32,133)
    for case in ("vertex_positions", "face_counts"):
        with open(f"{GOLDEN}/voropp/{case}/generators.txt") as fh:
            p = parse_generators(fh.read())
        assert p.shape == (100, 3)
        if case == "vertex_positions":
            assert np.array_equal(p[0], [-3.29554, 5.36459, -4.44451])
            assert np.array_equal(p[2], [0.258648, 6.78224, 2.2528])
