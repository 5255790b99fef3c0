"""Random patterns in the regex front end's syntax (tests only)."""


def random_pattern(rng) -> bytes:
    alpha = b"abcxyz019._-/"

    def lit():
        return bytes(alpha[int(i)] for i in rng.integers(0, len(alpha), int(rng.integers(1, 4)))).replace(
            b".", b"\\.")

    def piece(depth):
        k = rng.random()
        if k < 0.35:
            out = lit()
        elif k < 0.55:
            out = rng.choice([b"[a-c]", b"[^xyz]", b"\\d", b"\\w", b"[0-9_]", b"."]).item()
        elif k < 0.7 and depth < 2:
            out = b"(?:" + b"|".join(seq(depth + 1) for _ in range(int(rng.integers(2, 4)))) + b")"
        else:
            out = lit()
        q = rng.random()
        if q < 0.15:
            out += b"*"
        elif q < 0.25:
            out += b"+"
        elif q < 0.35:
            out += b"?"
        elif q < 0.42:
            a = int(rng.integers(0, 3))
            out += b"{%d,%d}" % (a, a + int(rng.integers(0, 3)))
        return out

    def seq(depth):
        parts = [piece(depth) for _ in range(int(rng.integers(1, 4)))]
        if len(parts) > 1 and rng.random() < 0.2:  # a boundary inside the pattern
            parts.insert(int(rng.integers(1, len(parts))), b"\\b")
        return b"".join(parts)

    body = seq(0)
    if rng.random() < 0.3:
        body = b"\\b" + body
    if rng.random() < 0.3:
        body = body + b"\\b"
    if rng.random() < 0.15:
        body = b"^" + body
    if rng.random() < 0.15:
        body = body + b"$"
    if rng.random() < 0.2:
        body = b"(?i)" + body
    return body
