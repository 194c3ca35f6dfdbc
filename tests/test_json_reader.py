"""The scene files' JSON reader against jsoncpp, the reader of the
reference's scene loader (std::istream >> Json::Value with jsoncpp's default
settings), both compiled side by side (oracle/_ref/libjson_check.so, built
from the reference tree).  For every text, both must accept or both refuse,
and accepted texts must give the same value tree (numbers as the exact
double and float that asDouble / asFloat return).  Texts: jsoncpp's own test
corpus, cases written for the jsoncpp extensions the reference's files may
use (comments, trailing commas, byte-order mark, lenient numbers), and
random mutations of valid documents.  Host only."""
import glob
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(not oracle.json_check_available(), reason="jsoncpp checker not built")
CORPUS = "/root/reference/ThirdParty/jsoncpp/test"


def _same(text):
    a = oracle.json_flatten(text, "ref")
    b = oracle.json_flatten(text, "pico")
    assert (a is None) == (b is None), (text[:200], a is None, b is None)
    assert a == b, text[:200]
    return a


@pytest.mark.skipif(not os.path.isdir(CORPUS), reason="jsoncpp's test corpus is not here")
def test_jsoncpp_corpus():
    files = sorted(glob.glob(os.path.join(CORPUS, "data", "*.json")) + glob.glob(os.path.join(CORPUS, "jsonchecker", "*.json")))
    assert len(files) > 80
    accepted = 0
    for f in files:
        with open(f, "rb") as fh:
            accepted += _same(fh.read()) is not None
    assert 40 < accepted < len(files)


CASES = [
    b'{"a": 1, "b": [1, 2, 3,], "c": {"d": null,},}',           # trailing commas
    b'// a scene\n{ /* block */ "a" : 1 // tail\n, "b": 2 }',        # comments
    b'{"a": 1 /* c */ "b": 2}', b'{"a": 1 /* c */ 5 "b": 2}',      # a token after a comment separates
    b'{"a" /* c */ : 1}', b'[ /* c */ ]', b'[1, /* c */ ]', b'[1 /* c */, 2]', b'[1 // c\r\n]',
    b'/*/ 1', b'/**/ [1]', b'/* open', b'// only', b'/x 1', b'[1, 2] trailing text {',
    b'\xef\xbb\xbf{"bom": true}', b'\xef\xbb\xbf', b'', b'   ',
] + [b"[" + n + b"]" for n in (
    b"-", b"+", b"1.", b"01", b"-0", b"-0.0", b"+1e5", b"1e", b"1e+", b".5", b"0x10", b"1.5e-3", b"-12345678901234567890",
    b"18446744073709551615", b"18446744073709551616", b"9223372036854775807", b"-9223372036854775808",
    b"-9223372036854775809", b"1e400", b"123456789012", b"16777217", b"9007199254740993", b"0.1", b"3.4028235e38",
    b"1e-320", b"-", b"--1", b"1-2", b"1.2.3", b"00.5e05")] + [
    b'[-Infinity]', b'[NaN]', b'[+]', b'[-]', b'[1.e5]', b'[1.5E+2]',
    b'["\\u00e9\\ud83d\\ude00\\udc00", "\\ud800\\u0041", "\\ud800x", "tab\there", "nl\nhere"]',
    b'["\\x"]', b'["\\u12"]', b'["\\', b'["open', b'{"a":1,"a":{"b":2},"a":[3]}', b'{"":0, "\\u0000":1}',
    b'{"k": tru}', b'{"k": nul}', b'{"k": falsey}', b'[true false]', b'{1: 2}', b"{'a': 1}", b'{"a" 1}',
    b'[' * 999 + b']' * 999, b'[' * 1001 + b']' * 1001,
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_extension_cases(k):
    _same(CASES[k])


def test_scene_like_document():
    text = b"""// the Cornell box
{
    "GLOBALS": { "Settings": { "SunDirection": [0.2, -1, 0.1], "SunColour": [5, 5, 4], } },
    /* meshes */
    "MESH": { "Box": { "Path": "box.obj" }, },
    "INSTANCE": {
        "b1": { "Asset": "Box", "BSRDF": "Dielectric", "Position": [-0.35, -0.7, 0.3], // left
                "Rotation": [0, 0.2588, 0, 0.9659], "Scale": [0.6, 0.6, 0.6] },
    },
}
"""
    tree = _same(text)
    assert tree is not None and "num" in tree


def test_mutations():
    """Random edits (deletions, insertions and replacements of structural
    characters) of valid documents."""
    rng = np.random.default_rng(7)
    seeds = [c for c in CASES[:3]] + [b'{"a": [1, -2.5e3, "x\\"y", {"b": [true, null]}], "c": "\\u00e9"}',
                                      b'[{"P": [0.1, 2, 3e-2], "S": "s"}, [], {}, 7]']
    alphabet = list(b'{}[],:"\\/*-+.eE0123456789 \n\tntfu') + [0xC3, 0xA9]
    accepted = 0
    for it in range(3000):
        t = bytearray(seeds[it % len(seeds)])
        for _ in range(int(rng.integers(1, 4))):
            op = int(rng.integers(0, 3))
            i = int(rng.integers(0, len(t) + 1))
            if op == 0 and len(t) > 1:
                del t[min(i, len(t) - 1)]
            elif op == 1:
                t.insert(i, alphabet[int(rng.integers(0, len(alphabet)))])
            elif len(t):
                t[min(i, len(t) - 1)] = alphabet[int(rng.integers(0, len(alphabet)))]
        accepted += _same(bytes(t)) is not None
    assert 300 < accepted < 2700, accepted
