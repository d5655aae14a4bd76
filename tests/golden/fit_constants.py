"""Fit constants of the reference's model functions, read mechanically from its source text.

Run in the build container (where the reference tree is mounted):
    python tests/golden/fit_constants.py
writes tests/golden/fit_constants.json: for each model function on the path, the file, the line range
of its body and the decimal literals of that body in source order (numbers with a '.' or an exponent;
bare integers are left out: they are array indices, loop bounds and small exponents such as pow(x, 2),
not fit constants).  Signs are operators, not part of a literal.

The functions (reference file:line):
  bispec          classed_version/bispectrum.cpp:166-236   BiHalofit (Takahashi et al. 2020, eqs. B4-B9)
  F2              classed_version/bispectrum.cpp:614-627   F2 with the d_n correction
  F2_tree         classed_version/bispectrum.cpp:629-633
  linear_pk_eh    classed_version/bispectrum.cpp:811-841   Eisenstein & Hu no-wiggle transfer function
  P_k_nonlinear   classed_version/bispectrum.cpp:288-358   revised Halofit (Takahashi et al. 2012)
  t17             cuda_version/bispectrum.cu:951-982      the Limber integrand with the T17 correction
  map2_t17        cuda_version/apertureStatistics.cu:39-46 the T17 factor of <Map^2>

The same extraction applied to the oracle's restatement (oracle/tp_oracle.cpp) must give the same
sequences: tests/test_fit_constants.py checks that, and evaluates an independent numpy restatement built
from these constants against the oracle's functions.
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"

# (name, file under the reference tree, regex of the function's first line)
REFERENCE_FUNCTIONS = [
    ("bispec", "classed_version/bispectrum.cpp", r"double\s+BispectrumCalculator::bispec\s*\(double k1"),
    ("F2", "classed_version/bispectrum.cpp", r"double\s+BispectrumCalculator::F2\s*\("),
    ("F2_tree", "classed_version/bispectrum.cpp", r"double\s+BispectrumCalculator::F2_tree\s*\("),
    ("linear_pk_eh", "classed_version/bispectrum.cpp", r"double\s+BispectrumCalculator::linear_pk_eh\s*\("),
    ("P_k_nonlinear", "classed_version/bispectrum.cpp", r"double\s+BispectrumCalculator::P_k_nonlinear\s*\("),
    ("t17", "cuda_version/bispectrum.cu",
     r"__host__ __device__ double limber_integrand_power_spectrum\(double ell, double z\)"),
]
# the T17 factor of <Map^2> sits inside a kernel: its lines are located by anchors
REFERENCE_SNIPPETS = [
    ("map2_t17", "cuda_version/apertureStatistics.cu", r"if \(T17_CORRECTION\)", r"dev_limber_integrand_power_spectrum"),
]

_LIT = re.compile(r"(?<![\w.])((?:\d+\.\d*|\.\d+)(?:[eE][+-]?\d+)?|\d+[eE][+-]?\d+)(?![\w.])")


def strip_comments(src):
    """Blanks // and /* */ comments (same length, newlines kept, so positions and lines stay put)."""
    blank = lambda m: re.sub(r"[^\n]", " ", m.group(0))  # noqa: E731
    return re.sub(r"//[^\n]*|/\*.*?\*/", blank, src, flags=re.S)


def function_body(src, first_line_regex):
    """(first line, last line, text) of the brace-delimited body that follows the first match (1-based)."""
    m = re.search(first_line_regex, src)
    if not m:
        raise ValueError("function not found: " + first_line_regex)
    i = src.index("{", m.end())
    depth = 0
    for j in range(i, len(src)):
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
            if depth == 0:
                return src.count("\n", 0, m.start()) + 1, src.count("\n", 0, j) + 1, src[i:j + 1]
    raise ValueError("unbalanced braces after " + first_line_regex)


def snippet(src, start_regex, end_regex):
    """(first line, last line, text) from the line of start_regex to the next line matching end_regex."""
    m = re.search(start_regex, src)
    if not m:
        raise ValueError("snippet not found: " + start_regex)
    e = re.compile(end_regex).search(src, m.end())
    if not e:
        raise ValueError("snippet end not found: " + end_regex)
    return src.count("\n", 0, m.start()) + 1, src.count("\n", 0, e.end()) + 1, src[m.start():e.end()]


def decimal_literals(text):
    """The decimal literals of a code fragment (comments removed) in order, as strings."""
    return _LIT.findall(strip_comments(text))


def extract_reference(root=REF):
    out = {}
    for name, rel, rx in REFERENCE_FUNCTIONS:
        src = strip_comments(open(os.path.join(root, rel)).read())
        a, b, body = function_body(src, rx)
        out[name] = {"file": rel, "lines": [a, b], "literals": decimal_literals(body)}
    for name, rel, rs, re_end in REFERENCE_SNIPPETS:
        src = strip_comments(open(os.path.join(root, rel)).read())
        a, b, body = snippet(src, rs, re_end)
        out[name] = {"file": rel, "lines": [a, b], "literals": decimal_literals(body)}
    return out


if __name__ == "__main__":
    fx = extract_reference()
    json.dump(fx, open(os.path.join(HERE, "fit_constants.json"), "w"), indent=1)
    for k, v in fx.items():
        print("%-14s %s:%d-%d  %d literals" % (k, v["file"], v["lines"][0], v["lines"][1], len(v["literals"])))
