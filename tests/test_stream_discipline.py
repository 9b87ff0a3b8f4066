"""Every copy and fill of the library goes through a stream the library
owns (its non-blocking stream, or a stream it created for a helper task),
never through the null stream: a non-blocking stream does not wait for the
null stream, and a fill on it once landed after the NaN-sentinel fill that
followed it.  A source check, run on the host."""
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = sorted(glob.glob(os.path.join(ROOT, "ib3d-gpu_amd", "csrc", "*.hip")) +
                 glob.glob(os.path.join(ROOT, "ib3d-gpu_amd", "csrc", "*.h")))


def _calls(text, name):
    """The argument text of every call of `name` (balanced parentheses)."""
    out = []
    for m in re.finditer(r"\b" + name + r"\s*\(", text):
        depth, i = 1, m.end()
        while depth and i < len(text):
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        out.append(text[m.end():i - 1])
    return out


def _strip_comments(t):
    t = re.sub(r"/\*.*?\*/", "", t, flags=re.S)
    return re.sub(r"//[^\n]*", "", t)


def test_no_null_stream_copies_or_fills():
    assert SOURCES
    for path in SOURCES:
        text = _strip_comments(open(path).read())
        for bad in ("hipMemcpy", "hipMemset", "hipMemcpy2D", "hipMemcpyToSymbol", "hipDeviceSynchronize"):
            assert not _calls(text, bad), (os.path.basename(path), bad)


def test_async_copies_and_fills_name_a_stream():
    streams = {"c.stream", "ctx().stream", "s", "stream"}
    n = 0
    for path in SOURCES:
        text = _strip_comments(open(path).read())
        for name in ("hipMemcpyAsync", "hipMemsetAsync"):
            for args in _calls(text, name):
                last = args.rsplit(",", 1)[-1].strip()
                assert last in streams, (os.path.basename(path), name, args)
                n += 1
    assert n >= 20
