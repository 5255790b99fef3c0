"""Seeded synthetic rule sets and HTTP-like inputs for the scan engine.

The reference consumes pre-compiled ``.nfa`` files only (its parser is
src/transition_graph.cpp:82-294) and ships no regex compiler, so the
workloads are generated here:

* a rule grammar of literal runs (4-12 bytes over ``[a-z0-9/=&?.%_-]``),
  character classes with bounded repetition ``{1,8}``, ``.*`` gaps and 2-3-way
  alternation, with optional case-insensitive literals;
* a Glushkov (position) automaton per rule: one state per position, every
  edge into a position labelled with that position's byte class, so the NFA
  is epsilon-free;
* the CRS file convention (data/crs/rule1.nfa:3-4,21-22): one shared
  initial state with a ``0|255`` self-loop (unanchored start) and a ``0|255``
  self-loop on every accepting position.  Acceptance at the end of a string
  is then exactly ``re.search(rule, string)``, which the tests use as an
  independent check of the oracle.

Rule ids are 1-based in generation order, like the merge numbering of
script/script.py:111-119.
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import List, Optional, Sequence, Tuple

import numpy as np

ALPHABET = b"abcdefghijklmnopqrstuvwxyz0123456789/=&?.%_-"
ALL_BYTES = frozenset(range(256))

# Byte classes used by class pieces.
_CLASSES = [
    frozenset(b"abcdefghijklmnopqrstuvwxyz"),
    frozenset(b"0123456789"),
    frozenset(b"abcdefghijklmnopqrstuvwxyz0123456789"),
    frozenset(b"0123456789abcdef"),
    frozenset(ALPHABET) - frozenset(b"&=?"),
    frozenset(b"abcdefghijklmnopqrstuvwxyz0123456789_-"),
    frozenset(b"%0123456789ABCDEF"),
]

# Tokens shared by rules and inputs so that rule prefixes occur in traffic.
VOCAB = [
    b"index", b"user", b"id=", b"session", b"search", b"page", b"/api/", b"login", b"token", b"admin",
    b"query", b"item", b"cart", b"img", b"static", b"v1", b"v2", b"lang=", b"sort=", b"ref=",
    b"uid=", b"name=", b"type=", b".php", b".js", b"list", b"view", b"data", b"file", b"path",
]


# ---------------------------------------------------------------- AST
@dataclasses.dataclass(frozen=True)
class Sym:
    syms: frozenset


@dataclasses.dataclass(frozen=True)
class Cat:
    items: tuple


@dataclasses.dataclass(frozen=True)
class Alt:
    items: tuple


@dataclasses.dataclass(frozen=True)
class Opt:
    item: object


@dataclasses.dataclass(frozen=True)
class Star:
    item: object


def _lit(s: bytes, icase: bool = False) -> Cat:
    out = []
    for ch in s:
        if icase and chr(ch).isalpha():
            out.append(Sym(frozenset({ord(chr(ch).lower()), ord(chr(ch).upper())})))
        else:
            out.append(Sym(frozenset({ch})))
    return Cat(tuple(out))


def _rep(node, lo: int, hi: int):
    """x{lo,hi} unrolled: lo mandatory copies then nested optionals."""
    tail = None
    for _ in range(hi - lo):
        tail = Opt(Cat((node,) if tail is None else (node, tail)))
    items = [node] * lo + ([tail] if tail is not None else [])
    return Cat(tuple(items))


# ---------------------------------------------------------------- regex text
def _esc_byte(b: int) -> bytes:
    return b"\\x%02x" % b


def _class_src(syms: frozenset) -> bytes:
    if len(syms) == 1:
        return _esc_byte(next(iter(syms)))
    if syms == ALL_BYTES:
        return b"."
    parts, run = [], None
    for b in sorted(syms):
        if run and b == run[1] + 1:
            run[1] = b
        else:
            if run:
                parts.append(run)
            run = [b, b]
    parts.append(run)
    body = b"".join(_esc_byte(a) if a == z else _esc_byte(a) + b"-" + _esc_byte(z) for a, z in parts)
    return b"[" + body + b"]"


def regex_source(node) -> bytes:
    """Python ``re`` (bytes, DOTALL) source equivalent to the AST."""
    if isinstance(node, Sym):
        return _class_src(node.syms)
    if isinstance(node, Cat):
        return b"".join(regex_source(x) for x in node.items)
    if isinstance(node, Alt):
        return b"(?:" + b"|".join(regex_source(x) for x in node.items) + b")"
    if isinstance(node, Opt):
        return b"(?:" + regex_source(node.item) + b")?"
    if isinstance(node, Star):
        return b"(?:" + regex_source(node.item) + b")*"
    raise TypeError(node)


def compile_regex(node) -> "re.Pattern":
    return re.compile(regex_source(node), re.DOTALL)


# ---------------------------------------------------------------- Glushkov
class _Glushkov:
    def __init__(self):
        self.labels: List[frozenset] = []
        self.follow: List[set] = []

    def build(self, node) -> Tuple[bool, set, set]:
        if isinstance(node, Sym):
            p = len(self.labels)
            self.labels.append(node.syms)
            self.follow.append(set())
            return False, {p}, {p}
        if isinstance(node, Cat):
            nullable, first, last = True, set(), set()
            for it in node.items:
                n2, f2, l2 = self.build(it)
                for p in last:
                    self.follow[p] |= f2
                if nullable:
                    first = first | f2
                last = (last | l2) if n2 else set(l2)
                nullable = nullable and n2
            return nullable, first, last
        if isinstance(node, Alt):
            nullable, first, last = False, set(), set()
            for it in node.items:
                n2, f2, l2 = self.build(it)
                nullable, first, last = nullable or n2, first | f2, last | l2
            return nullable, first, last
        if isinstance(node, Opt):
            _, f, l = self.build(node.item)
            return True, f, l
        if isinstance(node, Star):
            _, f, l = self.build(node.item)
            for p in l:
                self.follow[p] |= f
            return True, f, l
        raise TypeError(node)


def _atoms(syms: frozenset) -> str:
    out, run = [], None
    for b in sorted(syms):
        if run and b == run[1] + 1:
            run[1] = b
        else:
            if run:
                out.append(run)
            run = [b, b]
    out.append(run)
    return " ".join(str(a) if a == z else f"{a}|{z}" for a, z in out)


@dataclasses.dataclass
class RuleSet:
    rules: List[object]            # AST per rule (rule id = index + 1)
    nfa_text: str                  # merged .nfa file contents
    state_count: int
    gap_rules: int

    def write(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.nfa_text)

    def patterns(self):
        return [compile_regex(r) for r in self.rules]


def rules_to_nfa(rules: Sequence[object], header: str = "") -> Tuple[str, int]:
    """Merged .nfa text for ``rules`` with one shared initial state 0."""
    lines = ["# NFA" + (f" {header}" if header else "")]
    body = []
    next_id = 1
    body.append("0 : initial")
    body.append("0 -> 0 : 0|255")
    for rid, r in enumerate(rules, start=1):
        g = _Glushkov()
        nullable, first, last = g.build(r)
        if nullable:
            raise ValueError(f"rule {rid} is nullable")
        base = next_id
        gid = lambda p: base + p  # noqa: E731
        for q in sorted(first):
            body.append(f"0 -> {gid(q)} : {_atoms(g.labels[q])}")
        for p in range(len(g.labels)):
            for q in sorted(g.follow[p]):
                body.append(f"{gid(p)} -> {gid(q)} : {_atoms(g.labels[q])}")
        for p in sorted(last):
            body.append(f"{gid(p)} : accepting {rid}")
            body.append(f"{gid(p)} -> {gid(p)} : 0|255")
        next_id += len(g.labels)
    lines.append(str(next_id))
    lines.extend(body)
    return "\n".join(lines) + "\n", next_id


# ---------------------------------------------------------------- rule grammar
@dataclasses.dataclass
class GrammarConfig:
    n_rules: int
    icase_frac: float = 0.0      # fraction of literal runs made case-insensitive
    gap_rule_frac: float = 0.1   # fraction of rules containing a '.*' gap
    vocab_prefix_frac: float = 0.3  # literal runs starting with a traffic token
    pieces: Tuple[int, int] = (2, 4)
    # dense shapes: gap rules open on a traffic token plus one byte (the '.*'
    # right after it) and continue with a run of the burst class; a fraction
    # of the other rules start with such a run
    gap_after_token: bool = False
    class_lead_frac: float = 0.0
    gap_class_max: int = 8  # repetition bound of the class run after a gap


def _rand_literal(rng, lo=4, hi=12, vocab_frac=0.0) -> bytes:
    n = int(rng.integers(lo, hi + 1))
    if rng.random() < vocab_frac:
        tok = VOCAB[int(rng.integers(len(VOCAB)))]
        rest = bytes(ALPHABET[int(i)] for i in rng.integers(0, len(ALPHABET), size=max(0, n - len(tok))))
        return (tok + rest)[:max(n, len(tok))]
    return bytes(ALPHABET[int(i)] for i in rng.integers(0, len(ALPHABET), size=n))


def _dense_rule(rng, cfg: GrammarConfig, with_gap: bool):
    burst = Sym(BURST_CLASS)
    if with_gap:
        # token + one more byte, '.*', a burst-class run, a literal: the gap
        # opens on the (uncommon) pair, and from then on every burst-class
        # byte re-enters the run
        tok = VOCAB[int(rng.integers(len(VOCAB)))] + bytes([ALPHABET[int(rng.integers(36))]])
        items = [_lit(tok), Star(Sym(ALL_BYTES)), _rep(burst, 1, cfg.gap_class_max),
                 _lit(_rand_literal(rng), rng.random() < cfg.icase_frac)]
        return Cat(tuple(items))
    # a burst-class run, then literal pieces
    items = [_rep(burst, 1, cfg.gap_class_max), _lit(_rand_literal(rng, 4, 8), rng.random() < cfg.icase_frac)]
    for _ in range(int(rng.integers(cfg.pieces[0], cfg.pieces[1] + 1)) - 2):
        items.append(_lit(_rand_literal(rng), rng.random() < cfg.icase_frac))
    return Cat(tuple(items))


def random_rule(rng, cfg: GrammarConfig, with_gap: bool):
    # (no draw when there are no class-led rules: the other grammars keep
    # their random streams)
    if (with_gap and cfg.gap_after_token) or (not with_gap and cfg.class_lead_frac > 0
                                              and rng.random() < cfg.class_lead_frac):
        return _dense_rule(rng, cfg, with_gap)
    npieces = int(rng.integers(cfg.pieces[0], cfg.pieces[1] + 1))
    items = [_lit(_rand_literal(rng, vocab_frac=cfg.vocab_prefix_frac), rng.random() < cfg.icase_frac)]
    gap_at = int(rng.integers(1, npieces)) if with_gap and npieces >= 2 else -1
    for k in range(1, npieces):
        if k == gap_at:
            items.append(Star(Sym(ALL_BYTES)))
            # after a gap: a class run or a literal, then a closing literal
            if rng.random() < 0.5:
                cls = _CLASSES[int(rng.integers(len(_CLASSES)))]
                items.append(_rep(Sym(cls), 1, int(rng.integers(2, 9))))
            items.append(_lit(_rand_literal(rng), rng.random() < cfg.icase_frac))
            continue
        kind = rng.random()
        if kind < 0.4:
            items.append(_lit(_rand_literal(rng), rng.random() < cfg.icase_frac))
        elif kind < 0.75:
            cls = _CLASSES[int(rng.integers(len(_CLASSES)))]
            items.append(_rep(Sym(cls), 1, int(rng.integers(1, 9))))
        else:
            nb = int(rng.integers(2, 4))
            items.append(Alt(tuple(_lit(_rand_literal(rng), rng.random() < cfg.icase_frac) for _ in range(nb))))
    # a rule never ends on a nullable tail: close with a literal when needed
    if isinstance(items[-1], Star):
        items.append(_lit(_rand_literal(rng)))
    return Cat(tuple(items))


def make_ruleset(cfg: GrammarConfig, seed: int) -> RuleSet:
    rng = np.random.default_rng(seed)
    rules, gaps = [], 0
    for _ in range(cfg.n_rules):
        with_gap = rng.random() < cfg.gap_rule_frac
        r = random_rule(rng, cfg, with_gap)
        gaps += int(with_gap)
        rules.append(r)
    text, S = rules_to_nfa(rules, header=f"synthetic seed={seed} rules={cfg.n_rules}")
    return RuleSet(rules=rules, nfa_text=text, state_count=S, gap_rules=gaps)


# ---------------------------------------------------------------- front-end patterns
def pattern_text(node, rng=None, boundary_frac: float = 0.0) -> bytes:
    """The rule as a pattern in the regex front end's syntax
    (admatch_amd/regex.py): dot-all, and with probability boundary_frac a
    \\b before a leading / after a trailing word character."""
    body = regex_source(node)
    if rng is not None and isinstance(node, Cat) and node.items:
        word = frozenset(b"abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789_")
        first, last = node.items[0], node.items[-1]
        if isinstance(first, Cat) and first.items and isinstance(first.items[0], Sym) and \
                first.items[0].syms <= word and rng.random() < boundary_frac:
            body = b"\\b" + body
        if isinstance(last, Cat) and last.items and isinstance(last.items[-1], Sym) and \
                last.items[-1].syms <= word and rng.random() < boundary_frac:
            body = body + b"\\b"
    return b"(?s)" + body


def frontend_patterns(n_rules: int, seed: int, boundary_frac: float = 0.3) -> List[bytes]:
    """Patterns of the medium grammar (20% case-insensitive literals) with
    word boundaries, for the regex front end."""
    rs = make_ruleset(dataclasses.replace(grammar_for("medium"), n_rules=n_rules), seed)
    rng = np.random.default_rng(seed + 1)
    return [pattern_text(r, rng, boundary_frac) for r in rs.rules]


# ---------------------------------------------------------------- sampling
def sample_match(node, rng, max_gap: int = 24) -> bytes:
    """A random member of the rule's language (used to plant matches)."""
    if isinstance(node, Sym):
        if node.syms == ALL_BYTES:
            return bytes([ALPHABET[int(rng.integers(len(ALPHABET)))]])
        s = sorted(node.syms)
        return bytes([s[int(rng.integers(len(s)))]])
    if isinstance(node, Cat):
        return b"".join(sample_match(x, rng, max_gap) for x in node.items)
    if isinstance(node, Alt):
        return sample_match(node.items[int(rng.integers(len(node.items)))], rng, max_gap)
    if isinstance(node, Opt):
        return sample_match(node.item, rng, max_gap) if rng.random() < 0.5 else b""
    if isinstance(node, Star):
        k = int(rng.integers(0, max_gap + 1))
        return b"".join(sample_match(node.item, rng, max_gap) for _ in range(k))
    raise TypeError(node)


# ---------------------------------------------------------------- inputs
_METHODS = [b"GET ", b"POST ", b"PUT ", b"HEAD "]


# Long runs of one class in traffic (session tokens, encoded blobs): the
# byte class of the dense workload's bursts -- uppercase letters that the
# request line's method and protocol names (GET, POST, PUT, HEAD, HTTP) do
# not use, so ordinary requests never enter a burst.
BURST_CLASS = frozenset(b"BCFIJKLMNQRVWXYZ")


def _corpus(rng, nbytes: int, burst_frac: float = 0.0) -> np.ndarray:
    """HTTP-request-like ASCII text: method, path, key=value query.  With
    burst_frac > 0, that fraction of requests carries one more query value:
    a token of 24-96 bytes of BURST_CLASS."""
    chunks, total = [], 0
    alpha = np.frombuffer(ALPHABET[:36], dtype=np.uint8)
    burst = np.frombuffer(bytes(sorted(BURST_CLASS)), dtype=np.uint8)
    while total < nbytes:
        parts = [_METHODS[int(rng.integers(len(_METHODS)))], b"/"]
        for _ in range(int(rng.integers(1, 5))):
            if rng.random() < 0.5:
                parts.append(VOCAB[int(rng.integers(len(VOCAB)))])
            else:
                parts.append(alpha[rng.integers(0, 36, size=int(rng.integers(2, 10)))].tobytes())
            parts.append(b"/")
        parts.append(b"?")
        for q in range(int(rng.integers(1, 8))):
            if q:
                parts.append(b"&")
            parts.append(alpha[rng.integers(0, 26, size=int(rng.integers(2, 8)))].tobytes())
            parts.append(b"=")
            v = alpha[rng.integers(0, 36, size=int(rng.integers(1, 16)))].tobytes()
            if rng.random() < 0.15:
                v += b"%2F" + alpha[rng.integers(0, 36, size=3)].tobytes()
            parts.append(v)
        if burst_frac > 0 and rng.random() < burst_frac:
            parts.append(b"&sid=" + burst[rng.integers(0, burst.size, size=int(rng.integers(24, 97)))].tobytes())
        parts.append(b" HTTP/1.1 ")
        s = b"".join(parts)
        chunks.append(s)
        total += len(s)
    return np.frombuffer(b"".join(chunks), dtype=np.uint8)


@dataclasses.dataclass
class Inputs:
    buf: np.ndarray       # uint8, all strings concatenated
    offsets: np.ndarray   # int64, n+1
    planted: np.ndarray   # bool, n

    @property
    def n(self) -> int:
        return self.offsets.size - 1

    def string(self, i: int) -> bytes:
        return self.buf[self.offsets[i]: self.offsets[i + 1]].tobytes()

    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets)


def make_lengths(rng, n: int, dist: str, lo: int = 200, hi: int = 1500,
                 median: float = 600.0, sigma: float = 0.8, max_len: int = 8192) -> np.ndarray:
    if dist == "uniform":
        return rng.integers(lo, hi + 1, size=n).astype(np.int64)
    if dist == "lognormal":
        x = rng.lognormal(mean=math.log(median), sigma=sigma, size=n)
        return np.clip(np.round(x), 1, max_len).astype(np.int64)
    raise ValueError(dist)


def make_inputs(rules: Optional[Sequence[object]], n: int, lengths: np.ndarray, planted_frac: float,
                seed: int, burst_frac: float = 0.0) -> Inputs:
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, dtype=np.int64)
    total = int(lengths.sum())
    # strings are windows of a shared request corpus (at most 4 MB of text)
    corpus = _corpus(rng, min(total, 4 << 20) + int(lengths.max()) + 4096, burst_frac)
    offsets = np.zeros(n + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(lengths)
    starts = rng.integers(0, corpus.size - int(lengths.max()) - 1, size=n)
    buf = np.empty(total, dtype=np.uint8)
    for i in range(n):
        buf[offsets[i]: offsets[i + 1]] = corpus[starts[i]: starts[i] + lengths[i]]
    planted = np.zeros(n, dtype=bool)
    if rules:
        k = int(round(planted_frac * n))
        for i in rng.choice(n, size=k, replace=False):
            L = int(lengths[i])
            for _ in range(8):
                r = rules[int(rng.integers(len(rules)))]
                m = sample_match(r, rng)
                if len(m) <= L:
                    pos = int(rng.integers(0, L - len(m) + 1))
                    buf[offsets[i] + pos: offsets[i] + pos + len(m)] = np.frombuffer(m, np.uint8)
                    planted[i] = True
                    break
    return Inputs(buf=buf, offsets=offsets, planted=planted)


# ---------------------------------------------------------------- workloads
@dataclasses.dataclass
class Workload:
    name: str
    ruleset: RuleSet
    inputs: Inputs


# (rules seed, inputs seed) per workload, recorded here so runs are repeatable.
SEEDS = {"small": (1, 101), "medium": (2, 102), "large": (3, 103), "train": (3, 104), "serve": (3, 105),
         "large_sparse": (3, 103)}


def grammar_for(name: str) -> GrammarConfig:
    if name == "small":      # 64 rules, ~2k states
        return GrammarConfig(n_rules=64, icase_frac=0.0, gap_rule_frac=0.15, pieces=(3, 4))
    if name == "medium":     # 1,000 rules, ~20k states, 20% case-insensitive literals
        return GrammarConfig(n_rules=1000, icase_frac=0.2, gap_rule_frac=0.15, pieces=(2, 3))
    if name in ("large", "train", "serve"):
        # 3,000 rules, ~70k states, 30% with '.*' gaps (opening on a traffic
        # token plus one byte, then a burst-class run {1,8}), 45% of the rest
        # led by a burst-class run {1,8}: a few dozen states are active in
        # ordinary text, a few thousand on the uppercase method and protocol
        # names, and more than 10% of the states inside long burst-class
        # tokens (make_workload's burst_frac)
        return GrammarConfig(n_rules=3000, icase_frac=0.2, gap_rule_frac=0.3, pieces=(2, 4), gap_after_token=True,
                             class_lead_frac=0.45)
    if name == "large_sparse":
        # the same sizes without the dense shapes: the active set stays
        # below 0.3% of the states
        return GrammarConfig(n_rules=3000, icase_frac=0.2, gap_rule_frac=0.3, pieces=(2, 4))
    raise ValueError(name)


# Fraction of requests carrying a long burst-class token, per workload.
BURST_FRAC = {"large": 0.1, "train": 0.1, "serve": 0.1}


def make_workload(name: str, n: Optional[int] = None, seed_offset: int = 0) -> Workload:
    rseed, iseed = SEEDS[name]
    rs = make_ruleset(grammar_for(name), rseed)
    rng = np.random.default_rng(iseed + seed_offset)
    if name == "small":
        n = n or 1024
        lengths = make_lengths(rng, n, "uniform", 200, 1500)
        frac = 0.01
    elif name == "medium":
        n = n or 16384
        lengths = make_lengths(rng, n, "lognormal", median=600, sigma=0.8, max_len=8192)
        frac = 0.005
    elif name == "train":
        n = n or 8192
        lengths = make_lengths(rng, n, "lognormal", median=1024, sigma=0.8, max_len=16384)
        frac = 0.002
    elif name == "serve":
        n = n or 65536
        lengths = make_lengths(rng, n, "lognormal", median=800, sigma=0.8, max_len=16384)
        frac = 0.002
    else:  # large, large_sparse
        n = n or 65536
        lengths = make_lengths(rng, n, "lognormal", median=1024, sigma=0.8, max_len=16384)
        frac = 0.002
    inputs = make_inputs(rs.rules, n, lengths, frac, iseed + 7919 * seed_offset, BURST_FRAC.get(name, 0.0))
    return Workload(name=name, ruleset=rs, inputs=inputs)
