"""Regex front end: a PCRE subset compiled to the reference's ``.nfa`` format.

    python -m admatch_amd.regex rules.txt -o rules.nfa      (one pattern per line)

The reference consumes pre-compiled ``.nfa`` files only (its loader is
src/transition_graph.cpp:82-294); this module produces them from patterns,
one rule per pattern (rule id = line number among the non-empty lines, as
script/script.py:111-119 numbers merged files).

Supported syntax (bytes; a pattern is matched with search semantics):
  literals and escapes  \\xHH \\n \\r \\t \\f \\v \\0 and \\ before any punctuation
  .                     any byte but \\n ((?s) makes it any byte)
  [...] [^...]          classes with ranges, escapes and \\d \\w \\s \\D \\W \\S inside
  \\d \\w \\s \\D \\W \\S     ASCII classes
  * + ? {m} {m,} {m,n}  greedy or lazy (the same language)
  ( ) (?: )             groups
  |                     alternation
  \\b                    word boundary ([A-Za-z0-9_] vs the rest, string ends
                        count as non-word), as the engine's symbol 256
  ^ (at the start)      anchors the match at the start of the string
  $ (at the end)        anchors the match at the end of the string (\\z)
  (?i) (?s) (?is)       leading flags: ASCII case-insensitive, dot-all

Construction: a Glushkov (position) automaton per rule, as in synth.py; a
``\\b`` is a position entered through a word-boundary transition (symbol 256),
so the engine's boundary pass (src/iNFA_kernel.cu:135-142) moves a rule
across it.  Rules share state 0 (initial, ``0|255`` self-loop: unanchored
start); a ``^`` rule gets its own initial state without the self-loop.
Accepting positions get a ``0|255`` self-loop (a match anywhere stays
accepted to the end of the string) unless the rule ends in ``$``.
"""
from __future__ import annotations

import argparse
import dataclasses
import sys
from typing import List, Sequence, Tuple

from .synth import ALL_BYTES, Alt, Cat, Opt, Star, Sym, _atoms, _rep

WORD = frozenset(b"abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789_")
DIGIT = frozenset(b"0123456789")
SPACE = frozenset(b" \t\n\r\f\v")
WB_SYMBOL = 256


@dataclasses.dataclass(frozen=True)
class Wb:
    """Zero-width word boundary."""


class RegexError(ValueError):
    pass


@dataclasses.dataclass
class Rule:
    node: object
    anchored_start: bool
    anchored_end: bool


_ESC = {ord("n"): 10, ord("r"): 13, ord("t"): 9, ord("f"): 12, ord("v"): 11, ord("0"): 0}
_CLASS_ESC = {ord("d"): DIGIT, ord("w"): WORD, ord("s"): SPACE,
              ord("D"): ALL_BYTES - DIGIT, ord("W"): ALL_BYTES - WORD, ord("S"): ALL_BYTES - SPACE}


class _Parser:
    def __init__(self, src: bytes):
        self.s = src
        self.i = 0
        self.icase = False
        self.dotall = False

    def peek(self) -> int:
        return self.s[self.i] if self.i < len(self.s) else -1

    def take(self) -> int:
        if self.i >= len(self.s):
            raise RegexError("unexpected end of pattern")
        c = self.s[self.i]
        self.i += 1
        return c

    def error(self, msg: str):
        raise RegexError(f"{msg} at offset {self.i} of {self.s!r}")

    # ------------------------------------------------------------ entry
    def parse(self) -> Rule:
        while self.s.startswith(b"(?", self.i) and self.i + 2 < len(self.s) and self.s[self.i + 2] in b"is":
            j = self.i + 2
            while j < len(self.s) and self.s[j] in b"is":
                j += 1
            if j >= len(self.s) or self.s[j] != ord(")"):
                break
            flags = self.s[self.i + 2:j]
            self.icase |= ord("i") in flags
            self.dotall |= ord("s") in flags
            self.i = j + 1
        start = self.peek() == ord("^")
        if start:
            self.i += 1
        end = False
        body = self.s
        if body.endswith(b"$") and not body.endswith(b"\\$"):
            end = True
            self.s = body[:-1]
        node = self.alternation()
        if self.i != len(self.s):
            self.error("unbalanced ')'")
        if (start or end) and isinstance(node, Alt):
            self.error("an anchored top-level alternation must be grouped")
        return Rule(node=node, anchored_start=start, anchored_end=end)

    # ------------------------------------------------------------ grammar
    def alternation(self):
        items = [self.concatenation()]
        while self.peek() == ord("|"):
            self.i += 1
            items.append(self.concatenation())
        return items[0] if len(items) == 1 else Alt(tuple(items))

    def concatenation(self):
        items = []
        while self.peek() not in (-1, ord("|"), ord(")")):
            items.append(self.repetition())
        return Cat(tuple(items))

    def repetition(self):
        atom = self.atom()
        while True:
            c = self.peek()
            if c in (ord("*"), ord("+"), ord("?")):
                self.i += 1
                lo, hi = {ord("*"): (0, None), ord("+"): (1, None), ord("?"): (0, 1)}[c]
            elif c == ord("{") and self._is_bound():
                lo, hi = self._bound()
            else:
                return atom
            if self.peek() == ord("?"):  # lazy: same language
                self.i += 1
            if isinstance(atom, Wb):
                self.error("quantified \\b")
            atom = self._quantify(atom, lo, hi)

    def _is_bound(self) -> bool:
        j = self.i + 1
        seen_digit = False
        while j < len(self.s) and chr(self.s[j]).isdigit():
            j += 1
            seen_digit = True
        if not seen_digit:
            return False
        if j < len(self.s) and self.s[j] == ord("}"):
            return True
        if j < len(self.s) and self.s[j] == ord(","):
            j += 1
            while j < len(self.s) and chr(self.s[j]).isdigit():
                j += 1
            return j < len(self.s) and self.s[j] == ord("}")
        return False

    def _bound(self) -> Tuple[int, object]:
        self.i += 1
        j = self.s.index(b"}", self.i)
        body = self.s[self.i:j].decode()
        self.i = j + 1
        if "," in body:
            a, b = body.split(",")
            lo, hi = int(a), (int(b) if b else None)
        else:
            lo = hi = int(body)
        if hi is not None and hi < lo:
            self.error("bad repetition bounds")
        if max(lo, hi or 0) > 1000:
            self.error("repetition bound above 1000")
        return lo, hi

    @staticmethod
    def _quantify(atom, lo: int, hi):
        if hi is None:  # x{lo,} = lo copies then x*
            return Cat(tuple([atom] * lo + [Star(atom)]))
        if lo == 0 and hi == 1:
            return Opt(atom)
        if lo == hi == 0:
            return Cat(())
        return _rep(atom, lo, hi)

    def atom(self):
        c = self.take()
        if c == ord("("):
            if self.s.startswith(b"?:", self.i):
                self.i += 2
            elif self.peek() == ord("?"):
                self.error("unsupported group syntax")
            node = self.alternation()
            if self.take() != ord(")"):
                self.error("missing ')'")
            return node
        if c == ord("["):
            return Sym(self._class())
        if c == ord("."):
            return Sym(ALL_BYTES if self.dotall else ALL_BYTES - {10})
        if c == ord("\\"):
            e = self.take()
            if e == ord("b"):
                return Wb()
            if e in _CLASS_ESC:
                return Sym(_CLASS_ESC[e])
            return self._sym(frozenset({self._escape(e)}))
        if c in (ord("*"), ord("+"), ord("?")) or (c == ord("{") and self._bound_at(self.i - 1)):
            self.error("nothing to repeat")
        if c in (ord("^"), ord("$")):
            self.error("anchors are supported only at the ends of a pattern")
        return self._sym(frozenset({c}))

    def _bound_at(self, i: int) -> bool:
        save = self.i
        self.i = i
        ok = self._is_bound()
        self.i = save
        return ok

    def _escape(self, e: int) -> int:
        if e == ord("x"):
            h = self.s[self.i:self.i + 2]
            if len(h) != 2:
                self.error("bad \\x escape")
            self.i += 2
            return int(h, 16)
        if e in _ESC:
            return _ESC[e]
        if chr(e).isalnum():
            self.error(f"unsupported escape \\{chr(e)}")
        return e

    def _class(self) -> frozenset:
        neg = self.peek() == ord("^")
        if neg:
            self.i += 1
        out = set()
        first = True
        while True:
            c = self.take()
            if c == ord("]") and not first:
                break
            first = False
            if c == ord("\\"):
                e = self.take()
                if e in _CLASS_ESC:
                    out |= _CLASS_ESC[e]
                    continue
                lo = self._escape(e)
            else:
                lo = c
            if self.peek() == ord("-") and self.i + 1 < len(self.s) and self.s[self.i + 1] != ord("]"):
                self.i += 1
                h = self.take()
                hi = self._escape(self.take()) if h == ord("\\") else h
                if hi < lo:
                    self.error("bad class range")
                out |= set(range(lo, hi + 1))
            else:
                out.add(lo)
        syms = self._fold(frozenset(out))  # case folding before negation
        if not syms and not neg:
            self.error("empty class")
        return (ALL_BYTES - syms) if neg else syms

    def _fold(self, syms: frozenset) -> frozenset:
        if not self.icase:
            return syms
        extra = set()
        for b in syms:
            ch = chr(b)
            if ch.isascii() and ch.isalpha():
                extra.add(ord(ch.lower()))
                extra.add(ord(ch.upper()))
        return syms | frozenset(extra)

    def _sym(self, syms: frozenset) -> Sym:
        return Sym(self._fold(syms))


def parse(pattern) -> Rule:
    src = pattern.encode() if isinstance(pattern, str) else bytes(pattern)
    return _Parser(src).parse()


# ---------------------------------------------------------------- Glushkov with \b
class _Glushkov:
    """Position automaton; a \\b is a position labelled with symbol 256."""

    def __init__(self):
        self.labels: List[frozenset] = []
        self.follow: List[set] = []

    def _pos(self, label: frozenset):
        p = len(self.labels)
        self.labels.append(label)
        self.follow.append(set())
        return False, {p}, {p}

    def build(self, node):
        if isinstance(node, Sym):
            return self._pos(node.syms)
        if isinstance(node, Wb):
            return self._pos(frozenset({WB_SYMBOL}))
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


def compile_rules(patterns: Sequence, header: str = "") -> Tuple[str, int]:
    """Merged .nfa text for ``patterns`` (rule id = index + 1) and its state count."""
    rules = [parse(p) for p in patterns]
    body = ["0 : initial", "0 -> 0 : 0|255"]
    nxt = 1
    for rid, r in enumerate(rules, start=1):
        g = _Glushkov()
        nullable, first, last = g.build(r.node)
        src0 = 0
        if r.anchored_start:
            src0 = nxt
            nxt += 1
            body.append(f"{src0} : initial")
        base = nxt
        nxt += len(g.labels)
        gid = lambda p: base + p  # noqa: E731
        for q in sorted(first):
            body.append(f"{src0} -> {gid(q)} : {_atoms(g.labels[q])}")
        for p in range(len(g.labels)):
            for q in sorted(g.follow[p]):
                body.append(f"{gid(p)} -> {gid(q)} : {_atoms(g.labels[q])}")
        # the boundary pass follows word-boundary transitions in (src, dst)
        # order within one pass: a chain entered through a back edge (a
        # repetition looping to a \b that another \b follows) would stop early
        wb = [p for p, lab in enumerate(g.labels) if WB_SYMBOL in lab]
        for q in wb:
            if any(r in wb for r in g.follow[q]) and any(q in g.follow[p] for p in range(q + 1, len(g.labels))):
                raise RegexError(f"rule {rid}: word boundaries chained through a repetition")
        accepting = [gid(p) for p in sorted(last)]
        for s in accepting:
            body.append(f"{s} : accepting {rid}")
            if not r.anchored_end:
                body.append(f"{s} -> {s} : 0|255")
        if nullable:
            # the empty match: with both anchors only the empty string has it
            # (an initial state that dies on any byte); otherwise every string
            # has it (an initial state that never dies)
            e = nxt
            nxt += 1
            body += [f"{e} : initial", f"{e} : accepting {rid}"]
            if not (r.anchored_start and r.anchored_end):
                body.append(f"{e} -> {e} : 0|255")
    lines = ["# NFA" + (f" {header}" if header else ""), str(nxt)] + body
    return "\n".join(lines) + "\n", nxt


def main(argv=None):
    ap = argparse.ArgumentParser(description="compile one regex per line to a merged .nfa file")
    ap.add_argument("rules")
    ap.add_argument("-o", "--out", required=True)
    a = ap.parse_args(argv)
    with open(a.rules, "rb") as f:
        pats = [line.rstrip(b"\r\n") for line in f if line.strip()]
    try:
        text, S = compile_rules(pats, header=f"{len(pats)} rules from {a.rules}")
    except RegexError as e:
        print(f"regex: {e}", file=sys.stderr)
        return 2
    with open(a.out, "w") as f:
        f.write(text)
    print(f"{len(pats)} rules, {S} states -> {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
