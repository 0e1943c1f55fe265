"""Error behaviour of the C ABI without a GPU: every entry point checks its
handles and arguments before any HIP call, returns RTPG_ERR_INVALID (no
exception crosses the ABI) and leaves a message in the thread's last-error
string, which the Python wrapper raises as RtpgError with the status."""
import ctypes as C

import pytest

from rtlib_amd import _lib

INVALID = -1

# (entry point, arguments): NULL handles or NULL required arrays
CASES = [
    ("rtpg_render", (None, None)),
    ("rtpg_render_debug", (None, None, 0, None, None)),
    ("rtpg_render_relvar", (None, None)),
    ("rtpg_render_moments", (None, None, None)),
    ("rtpg_render_stats", (None, None)),
    ("rtpg_accum_clear", (None,)),
    ("rtpg_accum_download", (None, None, None)),
    ("rtpg_trace_closest", (None, None, 0, None)),
    ("rtpg_trace_closest_noalloc", (None, None, 0, None)),
    ("rtpg_adjoint", (None, None, None, None, None)),
    ("rtpg_sdtree_reset", (None, 1, 1)),
    ("rtpg_sdtree_build", (None,)),
    ("rtpg_sdtree_clear", (None,)),
    ("rtpg_sdtree_sizes", (None, None)),
    ("rtpg_sdtree_ctrl_beg_trace", (None,)),
    ("rtpg_sdtree_ctrl_end_trace", (None, 1)),
    ("rtpg_mlp_reset", (None, 0)),
    ("rtpg_mlp_train_recorded", (None, 1, None, None)),
]


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_null_handles_are_refused_with_a_message(name, args):
    lib = _lib.load()
    rc = getattr(lib, name)(*args)
    assert rc == INVALID, (name, rc)
    msg = lib.rtpg_last_error().decode()
    assert msg, name


def test_wrapper_raises_with_status_and_message():
    with pytest.raises(_lib.RtpgError) as e:
        _lib.call("rtpg_sdtree_reset", None, 1, 1)
    assert e.value.code == INVALID and "null" in str(e.value).lower()


def test_last_error_is_per_thread():
    import threading

    lib = _lib.load()
    lib.rtpg_sdtree_build(None)
    mine = lib.rtpg_last_error().decode()
    seen = []

    def other():
        seen.append(lib.rtpg_last_error().decode())

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert mine and seen == [""]
