"""A device allocation that fails while a render grows its buffers returns an
error and leaves no buffer that a later call could use after it was freed.

MPT_FAIL_ALLOC_AFTER=k makes the library's (k+1)-th device allocation after
the variable is set fail once (hipErrorOutOfMemory). A small render first
sizes the context's streaming path state (or the scene's pass buffers); a
larger render then has to grow them and is made to fail at the first, an
early, a middle and the last buffer of the growth. After each failure, with
the variable unset, the small render and the large one must both succeed and
give the films of undisturbed renders (the small one reuses nothing freed;
the large one reallocates what the failure left unsized)."""
import os

import numpy as np
import pytest

from cadise_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cadise_amd.api import Context
    return Context(0)


def _fail_once(k, call):
    from cadise_amd.native import MptError
    os.environ["MPT_FAIL_ALLOC_AFTER"] = str(k)
    try:
        with pytest.raises(MptError) as e:
            call()
        return str(e.value)
    finally:
        del os.environ["MPT_FAIL_ALLOC_AFTER"]


def test_failed_stream_growth_then_renders_recover(ctx):
    """The context's streaming path state: 2^16 slots for the small render, then
    2^21 for the large one, whose growth reallocates 34 buffers (2 x 16 queue
    arrays with the paths' state, 2 root queues); it fails at the 1st, 2nd,
    18th and 34th."""
    from cadise_amd.api import Context, Scene
    small = scenes.cornell_box(width=32, height=32, spp=1, max_depth=5)
    big = scenes.cornell_box(width=256, height=256, spp=32, max_depth=5)
    fresh = Context(0)
    want_small = Scene.from_spec(fresh, small).render(seed=5)
    want_big = Scene.from_spec(fresh, big).render(seed=6)
    del fresh
    a, b = Scene.from_spec(ctx, small), Scene.from_spec(ctx, big)
    tol = dict(rtol=2e-5, atol=1e-6)
    for k in (0, 1, 17, 33):
        ctx.release_stream_buffers()
        np.testing.assert_allclose(a.render(seed=5), want_small, **tol)
        msg = _fail_once(k, lambda: b.render(seed=6))
        print(f"[alloc failure, streaming, k={k}] {msg}", flush=True)
        np.testing.assert_allclose(a.render(seed=5), want_small, **tol)
        np.testing.assert_allclose(b.render(seed=6), want_big, **tol)


def test_failed_pass_buffers_then_render_recovers(ctx):
    """A scene's pass-by-pass buffers (path records, 2 x 16 queue arrays, queue
    counts, work counters: 35 allocations on its first render): the first render
    of a new scene fails at the 1st, 2nd, 18th and 35th; the next render of the
    same scene succeeds with the undisturbed film."""
    from cadise_amd.api import Context, Scene
    big = scenes.cornell_box(width=256, height=256, spp=2, max_depth=5)
    fresh = Context(0)
    want = Scene.from_spec(fresh, big).render(seed=6, passwise=True)
    del fresh
    for k in (0, 1, 17, 34):
        b = Scene.from_spec(ctx, big)
        msg = _fail_once(k, lambda: b.render(seed=6, passwise=True))
        print(f"[alloc failure, pass by pass, k={k}] {msg}", flush=True)
        np.testing.assert_array_equal(b.render(seed=6, passwise=True), want)
