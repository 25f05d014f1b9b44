"""Small helpers shared by tests (no tests here)."""


def pairs_of(path):
    """(s, d) pairs of a bare .edges/.txt or .mtx file as the loader reads them:
    leading comments and the first other line skipped, extra columns ignored."""
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines) and lines[i][:1] in ("#", "%"):
        i += 1
    out = []
    for line in lines[i + 1:]:
        tok = line.split()
        if len(tok) >= 2:
            out.append((int(tok[0]), int(tok[1])))
    return out
