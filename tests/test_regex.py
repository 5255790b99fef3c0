"""Regex front end (admatch_amd/regex.py): compiled .nfa files scanned by the
oracle against Python's re.search on the equivalent pattern, plus known
answers for word boundaries, anchors, case folding and empty matches."""
import re

import numpy as np
import pytest

from admatch_amd import regex as R
from regex_gen import random_pattern

import oracle


def py_pattern(p: bytes) -> re.Pattern:
    """The Python equivalent: ^ -> \\A, $ -> \\Z (the front end's $ is \\z)."""
    body, flags = p, 0
    m = re.match(rb"^\(\?([is]+)\)", body)
    if m:
        flags |= (re.I if b"i" in m.group(1) else 0) | (re.S if b"s" in m.group(1) else 0)
        body = body[m.end():]
    if body.startswith(b"^"):
        body = b"\\A" + body[1:]
    if body.endswith(b"$") and not body.endswith(b"\\$"):
        body = body[:-1] + b"\\Z"
    return re.compile(body, flags)


def scan(tmp_path, patterns, strings):
    text, _ = R.compile_rules(patterns)
    path = tmp_path / "r.nfa"
    path.write_text(text)
    o = oracle.OracleNFA(str(path))
    return [o.scan(s) for s in strings]


def expected(patterns, strings):
    pats = [py_pattern(p) for p in patterns]
    return [[i + 1 for i, p in enumerate(pats) if p.search(s)] for s in strings]


KNOWN = [
    (b"foo\\b", [b"foo", b"foo bar", b"a.foo", b"foox", b"foo_", b""]),
    (b"\\bfoo", [b"foo", b"xfoo", b" foo", b"_foo", b"-foo-"]),
    (b"\\bselect\\b.*\\bfrom\\b", [b"select a from b", b"selectfrom", b"xselect a from b", b"select afrom b"]),
    (b"^ab+c$", [b"abc", b"abbbc", b"xabc", b"abcx", b"ac", b""]),
    (b"^GET /", [b"GET /x", b" GET /x", b"get /"]),
    (b"(?i)Content-Type:\\s*text/xml", [b"content-type: TEXT/XML", b"Content-Type:text/xml", b"ContentType: text/xml"]),
    (b"(?i)x[^a]y", [b"xAy", b"xay", b"xby", b"XBY"]),
    (b"a.b", [b"a\nb", b"axb"]),
    (b"(?s)a.b", [b"a\nb", b"axb"]),
    (b"[\\d]{3}-\\d{2,}", [b"123-45", b"12-345", b"123-4", b"9999-999"]),
    (b"(?:ab|cd){2}", [b"abcd", b"abab", b"acbd"]),
    (b"\\x41\\x2e", [b"A.", b"A,"]),
    (b"a*", [b"", b"b"]),
    (b"^a*$", [b"", b"aaa", b"ab"]),
    (b"^(?:x|y)?$", [b"", b"x", b"xy"]),
    (b"colou?r", [b"color", b"colour", b"colouur"]),
    (b"[a-c-]+z", [b"-z", b"abz", b"dz"]),
    (b"\\.php\\?id=\\d+", [b"/index.php?id=12", b"/index.php?id=", b"/indexphp?id=1"]),
]


def test_known_answers(tmp_path):
    pats = [p for p, _ in KNOWN]
    strings = [s for _, ss in KNOWN for s in ss]
    assert scan(tmp_path, pats, strings) == expected(pats, strings)


def test_errors():
    for bad in [b"(ab", b"ab)", b"*a", b"a{3,2}", b"[b-a]", b"\\q", b"(?<n>a)", b"a^b", b"^a|b", b"\\b*",
                b"(?:\\b\\bx)*"]:
        with pytest.raises(R.RegexError):
            R.compile_rules([bad])


@pytest.mark.parametrize("seed", range(30))
def test_random_patterns_vs_python_re(tmp_path, seed):
    rng = np.random.default_rng(seed)
    pats = []
    while len(pats) < 25:
        p = random_pattern(rng)
        try:
            R.compile_rules([p])
        except R.RegexError:
            continue
        pats.append(p)
    alpha = np.frombuffer(b"abcxyzABCXYZ019._-/ \n", dtype=np.uint8)
    strings = [alpha[rng.integers(0, alpha.size, int(rng.integers(0, 30)))].tobytes() for _ in range(150)]
    got = scan(tmp_path, pats, strings)
    exp = expected(pats, strings)
    for s, g, e in zip(strings, got, exp):
        assert g == e, (s, g, e, [pats[i - 1] for i in set(g) ^ set(e)])


def test_cli(tmp_path):
    src = tmp_path / "rules.txt"
    src.write_bytes(b"foo\\b\n\n^GET /\n")
    out = tmp_path / "rules.nfa"
    assert R.main([str(src), "-o", str(out)]) == 0
    o = oracle.OracleNFA(str(out))
    assert o.scan(b"GET /foo") == [1, 2]
