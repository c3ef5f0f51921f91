"""Allocation-free contact generation (``na_collide_noalloc``) through the C ABI.

The entry point keeps ``na_collide``'s staging area in a static device array instead of the
caller's scratch.  Its contacts must equal ``na_collide``'s bit for bit (integer and float arrays,
counts and per-world counts), on two calls in a row, the second with a smaller input, so that the
array's contents from the first call cannot leak into the second.  A request larger than the
static array is refused with an error.
"""

import ctypes as C
import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
na = importlib.import_module("newton-adaptive_amd")
_lib = importlib.import_module("newton-adaptive_amd._lib")
_native = importlib.import_module("newton-adaptive_amd.native")

FIELDS = ("rigid_contact_shape0", "rigid_contact_shape1", "rigid_contact_point0", "rigid_contact_point1",
          "rigid_contact_offset0", "rigid_contact_offset1", "rigid_contact_normal", "rigid_contact_margin0",
          "rigid_contact_margin1", "rigid_contact_count")


def _desc(contacts, world_counts, scratch):
    d = _lib.NaContactsDesc()
    d.capacity = contacts.rigid_contact_max
    for f in ("count", "shape0", "shape1", "point0", "point1", "offset0", "offset1", "normal", "margin0", "margin1"):
        setattr(d, f, getattr(contacts, "rigid_contact_" + f).data_ptr())
    d.world_count_out = world_counts.data_ptr()
    d.scratch = scratch.data_ptr() if scratch is not None else None
    return d


def _settled(worlds, seed, max_prims, frames):
    model = na.scenes.uneven_pile(worlds, seed=seed, device=DEV, max_prims=max_prims)
    state = model.state()
    solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128)
    for _ in range(frames):
        solver.step(state, state, model.control(), None, 0.01)
    torch.cuda.synchronize()
    return model, state


def _both(model, state, per_world):
    """Contacts of na_collide (caller scratch) and na_collide_noalloc (static scratch)."""
    L = _lib.lib()
    h = _native.model_handle(model)
    W = model.world_count
    out = []
    for variant in ("na_collide", "na_collide_noalloc"):
        contacts = na.Contacts(per_world * W, model.device)
        for f in FIELDS:  # poison every output so that stale or unwritten values show
            getattr(contacts, f).fill_(-7)
        wc = torch.full((W,), -7, dtype=torch.int32, device=DEV)
        scratch = None
        if variant == "na_collide":
            words = int(L.na_collide_scratch_words(h.handle, per_world))
            scratch = torch.full((words,), -7, dtype=torch.int32, device=DEV)
        d = _desc(contacts, wc, scratch)
        print(f"{variant}: worlds={W} per_world={per_world}", flush=True)
        _lib.check(getattr(L, variant)(h.handle, state.body_q.data_ptr(), C.byref(d), per_world, 0,
                                       _lib.stream_handle()), variant)
        torch.cuda.synchronize()
        n = int(contacts.rigid_contact_count.item())
        m = min(n, contacts.rigid_contact_max)
        res = {f: getattr(contacts, f)[:m].cpu().numpy() for f in FIELDS if f != "rigid_contact_count"}
        res["count"] = n
        res["world_counts"] = wc.cpu().numpy()
        out.append(res)
    return out


def test_noalloc_collide_equals_collide_on_two_calls():
    big = _settled(192, 3, 32, 12)
    small = _settled(24, 8, 10, 4)
    for (model, state), per_world in ((big, 128), (small, 64)):
        ref, got = _both(model, state, per_world)
        assert ref["count"] > 0 and ref["count"] == got["count"]
        assert np.array_equal(ref["world_counts"], got["world_counts"])
        for k in ref:
            if k in ("count", "world_counts"):
                continue
            assert ref[k].tobytes() == got[k].tobytes(), f"{k} differs (worlds={model.world_count})"


def test_noalloc_collide_refuses_sizes_beyond_the_static_scratch():
    model = na.scenes.box_drop(8, seed=1, device=DEV)
    h = _native.model_handle(model)
    words = int(_lib.lib().na_collide_scratch_words(h.handle, 128))
    assert words <= 12 * 1024 * 1024
    # a model with more world-slots than the array holds: 5000 worlds x 128 contacts x 20 words
    big = na.scenes.box_drop(5000, seed=1, device=DEV)
    contacts = na.Contacts(128, big.device)
    wc = torch.zeros(big.world_count, dtype=torch.int32, device=DEV)
    d = _desc(contacts, wc, None)
    with pytest.raises(RuntimeError, match="exceeds the static scratch"):
        _lib.check(_lib.lib().na_collide_noalloc(_native.model_handle(big).handle, big.body_q.data_ptr(),
                                                 C.byref(d), 128, 0, _lib.stream_handle()), "na_collide_noalloc")
