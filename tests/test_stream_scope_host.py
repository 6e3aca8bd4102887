"""The stream decorator of the library bindings (isogriddft_amd._lib.on_stream), on the CPU:
with a stream the wrapped body runs inside ``torch.cuda.stream(stream)`` and sees
``stream=None`` (so every launch inside it takes the current stream, now that stream);
without one it runs unchanged.  Positional, keyword and keyword-only stream arguments and
methods are covered; the stream context is recorded by a stand-in."""

import contextlib

import torch

from isogriddft_amd import _lib


def _recording(monkeypatch):
    entered = []

    @contextlib.contextmanager
    def fake_stream(s):
        entered.append(s)
        yield

    monkeypatch.setattr(torch.cuda, "stream", fake_stream)
    return entered


def test_positional_and_keyword_streams(monkeypatch):
    entered = _recording(monkeypatch)

    @_lib.on_stream
    def f(a, b=1, stream=None, c=2):
        return (a, b, stream, c)

    assert f(5) == (5, 1, None, 2) and entered == []
    assert f(5, 3, None, 4) == (5, 3, None, 4) and entered == []
    assert f(5, 3, "S1", 4) == (5, 3, None, 4) and entered == ["S1"]
    assert f(5, stream="S2", c=7) == (5, 1, None, 7) and entered == ["S1", "S2"]


def test_keyword_only_stream_and_methods(monkeypatch):
    entered = _recording(monkeypatch)

    @_lib.on_stream
    def g(a, *, tol=1.0, stream=None):
        return (a, tol, stream)

    class K:
        @_lib.on_stream
        def m(self, x, out=None, stream=None):
            return (self, x, out, stream)

    assert g(1, tol=2.0) == (1, 2.0, None) and entered == []
    assert g(1, stream="S3") == (1, 1.0, None) and entered == ["S3"]
    k = K()
    assert k.m(4, 5, "S4") == (k, 4, 5, None) and entered == ["S3", "S4"]
    assert k.m(4) == (k, 4, None, None) and entered == ["S3", "S4"]
