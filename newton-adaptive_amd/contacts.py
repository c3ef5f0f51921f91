"""``Contacts`` buffer and ``CollisionPipeline`` with the reference's surface.

``Contacts`` keeps the ``rigid_contact_*`` SoA names, shapes and dtypes of
``newton/_src/sim/contacts.py:137-345``; ``CollisionPipeline`` keeps
``CollisionPipeline(model, broad_phase=..., rigid_contact_max=...)``, ``.contacts()`` and
``.collide(state, contacts)`` of ``newton/_src/sim/collide.py:485-511,875,920``.  Contact
generation runs in ``libnadapt.so`` (one wavefront per world).  The generated order is already
run-to-run identical (worlds in order, then candidate pairs in sweep order, then manifold slot);
``deterministic=True`` (implied by contact matching, ``collide.py:618-620``) re-sorts it on the
device into the reference's canonical order, ascending ``make_contact_sort_key(shape0, shape1,
sub)`` (``contact_data.py:59-83``, ``contact_sort.py:194``), so the contact rows and match indices
are the ones the reference's deterministic pipeline reports.
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .native import model_handle

BROAD_PHASES = {"sap": 0, "nxn": 1, "explicit": 2}
MATCHING_MODES = ("disabled", "latest", "sticky")  # collide.py:572-575
SORT_KEY_SENTINEL = 0x7FFFFFFFFFFFFFFF  # contact_sort.py
MATCH_NOT_FOUND = -1  # contact_match.py:113
MATCH_BROKEN = -2  # contact_match.py:116


class Contacts:
    """Rigid contact SoA buffer (``contacts.py:91``).  With ``contact_matching`` it also holds
    ``rigid_contact_match_index`` and, with ``contact_report``, the new / broken index lists
    (``contacts.py:278-310``)."""

    def __init__(self, rigid_contact_max: int, device, contact_matching: bool = False, contact_report: bool = False):
        if contact_report and not contact_matching:
            raise ValueError("contact_report=True requires contact_matching=True")
        n = int(rigid_contact_max)
        d = torch.device(device)
        self.rigid_contact_max = n
        self.device = d
        self.rigid_contact_count = torch.zeros(1, dtype=torch.int32, device=d)
        self.rigid_contact_shape0 = torch.full((n,), -1, dtype=torch.int32, device=d)
        self.rigid_contact_shape1 = torch.full((n,), -1, dtype=torch.int32, device=d)
        self.rigid_contact_point0 = torch.zeros((n, 3), dtype=torch.float32, device=d)
        self.rigid_contact_point1 = torch.zeros((n, 3), dtype=torch.float32, device=d)
        self.rigid_contact_offset0 = torch.zeros((n, 3), dtype=torch.float32, device=d)
        self.rigid_contact_offset1 = torch.zeros((n, 3), dtype=torch.float32, device=d)
        self.rigid_contact_normal = torch.zeros((n, 3), dtype=torch.float32, device=d)
        self.rigid_contact_margin0 = torch.zeros(n, dtype=torch.float32, device=d)
        self.rigid_contact_margin1 = torch.zeros(n, dtype=torch.float32, device=d)
        self.rigid_contact_tids = torch.full((n,), -1, dtype=torch.int32, device=d)
        self.rigid_contact_world_count = None
        self.contact_matching = bool(contact_matching)
        self.contact_report = bool(contact_report)
        self.rigid_contact_match_index = torch.full((n,), -1, dtype=torch.int32, device=d) if contact_matching \
            else None
        if contact_report:
            self.rigid_contact_new_indices = torch.zeros(n, dtype=torch.int32, device=d)
            self.rigid_contact_new_count = torch.zeros(1, dtype=torch.int32, device=d)
            self.rigid_contact_broken_indices = torch.zeros(n, dtype=torch.int32, device=d)
            self.rigid_contact_broken_count = torch.zeros(1, dtype=torch.int32, device=d)
        else:
            self.rigid_contact_new_indices = self.rigid_contact_new_count = None
            self.rigid_contact_broken_indices = self.rigid_contact_broken_count = None

    def clear(self) -> None:
        self.rigid_contact_count.zero_()
        self.rigid_contact_shape0.fill_(-1)
        self.rigid_contact_shape1.fill_(-1)
        if self.rigid_contact_match_index is not None:
            self.rigid_contact_match_index.fill_(-1)


class CollisionPipeline:
    """AABB -> broad phase -> primitive narrow phase -> ``Contacts`` (``collide.py:464``).

    ``deterministic`` sorts the contacts into the reference's key order after generation;
    ``contact_matching`` (``"disabled"``, ``"latest"``, ``"sticky"``; ``True`` means
    ``"latest"``) matches every contact against the previous frame's and implies
    ``deterministic`` (``collide.py:572-620``)."""

    def __init__(self, model, *, broad_phase: str = "sap", rigid_contact_max: int | None = None,
                 rigid_contact_max_per_world: int = 128, deterministic: bool = False,
                 contact_matching="disabled", contact_report: bool = False,
                 contact_matching_pos_threshold: float = 0.0005, contact_matching_normal_dot_threshold: float = 0.995,
                 **_kw):
        if contact_matching is True:
            contact_matching = "latest"
        elif contact_matching is False or contact_matching is None:
            contact_matching = "disabled"
        if contact_matching not in MATCHING_MODES:
            raise ValueError(f"contact_matching must be one of {MATCHING_MODES}, got {contact_matching!r}")
        if contact_matching_pos_threshold < 0.0:
            raise ValueError(f"contact_matching_pos_threshold must be non-negative, got "
                             f"{contact_matching_pos_threshold}")
        if not -1.0 <= contact_matching_normal_dot_threshold <= 1.0:
            raise ValueError(f"contact_matching_normal_dot_threshold must be in [-1, 1], got "
                             f"{contact_matching_normal_dot_threshold}")
        if contact_report and contact_matching == "disabled":
            raise ValueError('contact_report=True requires contact_matching != "disabled"')
        if broad_phase not in BROAD_PHASES:
            raise ValueError(f"broad_phase must be one of {tuple(BROAD_PHASES)}, got {broad_phase!r}")
        if not 1 <= int(rigid_contact_max_per_world) <= 128:
            raise ValueError("rigid_contact_max_per_world must be in [1, 128]")
        self.model = model
        self.broad_phase = broad_phase
        self.per_world = int(rigid_contact_max_per_world)
        self.rigid_contact_max = int(rigid_contact_max) if rigid_contact_max is not None \
            else self.per_world * model.world_count
        self._handle = model_handle(model, with_pairs=broad_phase == "explicit")
        L = _lib.lib()
        words = int(L.na_collide_scratch_words(self._handle.handle, self.per_world))
        self._scratch = torch.zeros(words, dtype=torch.int32, device=model.device)
        self._world_counts = torch.zeros(model.world_count, dtype=torch.int32, device=model.device)
        # frame-to-frame matching state (ContactMatcher, collide.py:853-861)
        self.contact_matching_mode = contact_matching
        self.contact_matching = contact_matching != "disabled"
        self.contact_report = bool(contact_report)
        self.deterministic = bool(deterministic) or self.contact_matching
        self.pos_threshold = float(contact_matching_pos_threshold)
        self.normal_dot_threshold = float(contact_matching_normal_dot_threshold)
        n, d = self.rigid_contact_max, model.device
        i32 = dict(dtype=torch.int32, device=d)
        if self.deterministic:
            if model.shape_count > 0xFFFFF:
                raise ValueError("deterministic contact order supports at most 1048575 shapes (20-bit key fields)")
            nbytes = _lib.check(int(L.na_sort_contacts_workspace_bytes(n)), "na_sort_contacts_workspace_bytes")
            self._sort_workspace = torch.empty(nbytes, dtype=torch.uint8, device=d)
            self._sort_key_array = torch.full((n,), SORT_KEY_SENTINEL, dtype=torch.int64, device=d)
            self._sort_rank = torch.arange(n, **i32)  # generated position -> sorted position, this frame
            self._prev_rank = torch.arange(n, **i32)  # the same for the previous frame
        if self.contact_matching:
            W = model.world_count
            # previous frame, in generated (world-major) order
            self._prev_s0 = torch.full((n,), -1, **i32)
            self._prev_s1 = torch.full((n,), -1, **i32)
            self._prev_mid = torch.zeros((n, 3), dtype=torch.float32, device=d)
            self._prev_n = torch.zeros((n, 3), dtype=torch.float32, device=d)
            self._prev_off = torch.zeros(W, **i32)
            self._prev_cnt = torch.zeros(W, **i32)
            self._prev_total = torch.zeros(1, **i32)
            self._prev_matched = torch.zeros(n, **i32)
            self._cur_mid = torch.zeros((n, 3), dtype=torch.float32, device=d)
            if contact_matching == "sticky":
                self._prev_geom = [torch.zeros((n, 3), dtype=torch.float32, device=d) for _ in range(4)]

    def reset_matching(self) -> None:
        """Forget the previous frame (the next collide reports every contact as new)."""
        if self.contact_matching:
            self._prev_cnt.zero_()
            self._prev_total.zero_()

    def contacts(self) -> Contacts:
        return Contacts(self.rigid_contact_max, self.model.device, contact_matching=self.contact_matching,
                        contact_report=self.contact_report)

    def collide(self, state, contacts: Contacts, stream=None) -> Contacts:
        if contacts.rigid_contact_max < 1:
            raise ValueError("contacts buffer has no capacity")
        if self.deterministic and contacts.rigid_contact_max != self.rigid_contact_max:
            raise ValueError(f"contacts capacity ({contacts.rigid_contact_max}) does not match the pipeline's "
                             f"deterministic sort buffer size ({self.rigid_contact_max}); use pipeline.contacts()")
        d = _lib.NaContactsDesc()
        d.capacity = contacts.rigid_contact_max
        d.count = contacts.rigid_contact_count.data_ptr()
        d.shape0 = contacts.rigid_contact_shape0.data_ptr()
        d.shape1 = contacts.rigid_contact_shape1.data_ptr()
        d.point0 = contacts.rigid_contact_point0.data_ptr()
        d.point1 = contacts.rigid_contact_point1.data_ptr()
        d.offset0 = contacts.rigid_contact_offset0.data_ptr()
        d.offset1 = contacts.rigid_contact_offset1.data_ptr()
        d.normal = contacts.rigid_contact_normal.data_ptr()
        d.margin0 = contacts.rigid_contact_margin0.data_ptr()
        d.margin1 = contacts.rigid_contact_margin1.data_ptr()
        d.world_count_out = self._world_counts.data_ptr()
        d.scratch = self._scratch.data_ptr()
        body_q = state.body_q.contiguous()
        L = _lib.lib()
        _lib.check(L.na_collide(self._handle.handle, body_q.data_ptr(), C.byref(d), self.per_world,
                                BROAD_PHASES[self.broad_phase], _lib.stream_handle(stream)), "na_collide")
        contacts.rigid_contact_world_count = self._world_counts
        if self.contact_matching:
            self._match(body_q, d, contacts, stream)
        if self.deterministic:
            mi = contacts.rigid_contact_match_index if self.contact_matching else None
            _lib.check(L.na_sort_contacts(
                C.byref(d), self.model.shape_count, contacts.rigid_contact_tids.data_ptr(),
                mi.data_ptr() if mi is not None else None,
                self._prev_rank.data_ptr() if mi is not None else None, self.rigid_contact_max,
                self._sort_key_array.data_ptr(), self._sort_rank.data_ptr(), self._sort_workspace.data_ptr(),
                self._sort_workspace.numel(), _lib.stream_handle(stream)), "na_sort_contacts")
        if self.contact_report:
            self._report(contacts)
        if self.deterministic:
            self._prev_rank, self._sort_rank = self._sort_rank, self._prev_rank
        return contacts

    def _match(self, body_q, d, contacts, stream):
        """Match the generated contacts against the previous frame's (indices in the previous
        frame's generated order; the deterministic sort maps them to its sorted order), replay the
        previous geometry in sticky mode, and keep this frame as the previous one."""
        if contacts.rigid_contact_match_index is None:
            raise ValueError("contacts were allocated without contact_matching")
        W = self.model.world_count
        n = self.rigid_contact_max
        cur_off = self._scratch[W * self.per_world * 20: W * self.per_world * 20 + W]
        contacts.rigid_contact_match_index.fill_(MATCH_NOT_FOUND)
        self._prev_matched.zero_()
        geom = [g.data_ptr() for g in self._prev_geom] if self.contact_matching_mode == "sticky" else [None] * 4
        _lib.check(_lib.lib().na_match_contacts(
            self._handle.handle, body_q.data_ptr(), C.byref(d), self.per_world, cur_off.data_ptr(),
            self._prev_s0.data_ptr(), self._prev_s1.data_ptr(), self._prev_mid.data_ptr(), self._prev_n.data_ptr(),
            self._prev_off.data_ptr(), self._prev_cnt.data_ptr(), n, self.pos_threshold, self.normal_dot_threshold,
            contacts.rigid_contact_match_index.data_ptr(), self._cur_mid.data_ptr(), self._prev_matched.data_ptr(),
            *geom, _lib.stream_handle(stream)), "na_match_contacts")
        self._report_prev_total = self._prev_total.clone()
        self._report_prev_matched = self._prev_matched
        # this frame becomes the previous one (generated order, after the sticky replay)
        self._prev_s0.copy_(contacts.rigid_contact_shape0)
        self._prev_s1.copy_(contacts.rigid_contact_shape1)
        self._prev_mid.copy_(self._cur_mid)
        self._prev_n.copy_(contacts.rigid_contact_normal)
        if self.contact_matching_mode == "sticky":
            for dst, src in zip(self._prev_geom, (contacts.rigid_contact_point0, contacts.rigid_contact_point1,
                                                  contacts.rigid_contact_offset0, contacts.rigid_contact_offset1)):
                dst.copy_(src)
        self._prev_off.copy_(cur_off)
        self._prev_cnt.copy_(self._world_counts)
        self._prev_total.copy_(torch.clamp(contacts.rigid_contact_count, max=n))

    def _report(self, contacts):
        """New / broken contact lists (``ContactMatcher.build_report``): new = current rows whose
        match index is negative; broken = previous-frame rows (sorted order) no current contact
        matched, ascending."""
        n = self.rigid_contact_max
        total = int(min(int(contacts.rigid_contact_count.item()), n))
        prev_total = int(self._report_prev_total.item())
        new = torch.nonzero(contacts.rigid_contact_match_index[:total] < 0).flatten().to(torch.int32)
        broken = torch.nonzero(self._report_prev_matched[:prev_total] == 0).flatten()
        if self.deterministic:
            broken = torch.sort(self._prev_rank[broken]).values
        broken = broken.to(torch.int32)
        contacts.rigid_contact_new_indices[:len(new)] = new
        contacts.rigid_contact_new_count.fill_(len(new))
        contacts.rigid_contact_broken_indices[:len(broken)] = broken
        contacts.rigid_contact_broken_count.fill_(len(broken))

    def debug_pairs(self, state, max_pairs: int = 1024, stream=None):
        """AABBs and candidate pairs (for pair-set parity checks): (lo [S,3], hi [S,3],
        pairs [W, max_pairs, 2], counts [W])."""
        m = self.model
        lo = torch.zeros((m.shape_count, 3), dtype=torch.float32, device=m.device)
        hi = torch.zeros_like(lo)
        pairs = torch.full((m.world_count, max_pairs, 2), -1, dtype=torch.int32, device=m.device)
        cnt = torch.zeros(m.world_count, dtype=torch.int32, device=m.device)
        _lib.check(_lib.lib().na_collide_debug(self._handle.handle, state.body_q.contiguous().data_ptr(),
                                               lo.data_ptr(), hi.data_ptr(), pairs.data_ptr(), cnt.data_ptr(),
                                               max_pairs, BROAD_PHASES[self.broad_phase],
                                               _lib.stream_handle(stream)), "na_collide_debug")
        return lo, hi, pairs, cnt
