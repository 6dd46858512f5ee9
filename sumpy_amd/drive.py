"""``drive_fmm``: the FMM pass over a wrangler.

Same call contract and stage order as ``boxtree.fmm.drive_fmm`` (the caller
of ``SumpyExpansionWrangler`` in ``sumpy/test/test_fmm.py:280-282``; the stage
list is that of the wrangler interface, ``sumpy/fmm.py:543-1086``).  boxtree
is not part of this build.  Level starts are handed to the wrangler from the
traversal's host copies, so a pass enqueues all its kernels without a
device-to-host synchronisation.
"""

from __future__ import annotations

import numpy as np


class _PhaseTimer:
    """Records HIP events around each stage on the current stream."""

    def __init__(self, enabled):
        self.enabled = enabled
        self.events = []

    def mark(self, name):
        if not self.enabled:
            return
        import torch
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.events.append((name, ev))

    def result(self):
        import torch
        torch.cuda.synchronize()
        out = {}
        for (n0, e0), (_, e1) in zip(self.events[:-1], self.events[1:]):
            out[n0] = out.get(n0, 0.0) + e0.elapsed_time(e1)
        return out


def drive_fmm(actx, wrangler, src_weight_vecs, timing_data=None):
    """Evaluate the potentials of ``src_weight_vecs`` through ``wrangler``.

    Returns the potentials in user target order.  If *timing_data* is a dict,
    it receives per-stage GPU times in milliseconds.
    """
    trav = wrangler.traversal
    h = trav.host if hasattr(trav, "host") else {}

    def host(name):
        return h[name] if name in h else getattr(trav, name)

    timer = _PhaseTimer(timing_data is not None)

    timer.mark("reorder_sources")
    src_weight_vecs = [wrangler.reorder_sources(w) for w in src_weight_vecs]
    src_weight_vecs = wrangler.distribute_source_weights(actx, src_weight_vecs, None)

    timer.mark("form_multipoles")
    mpole_exps = wrangler.form_multipoles(
        actx, host("level_start_source_box_nrs"), trav.source_boxes, src_weight_vecs)

    timer.mark("coarsen_multipoles")
    mpole_exps = wrangler.coarsen_multipoles(
        actx, host("level_start_source_parent_box_nrs"), trav.source_parent_boxes,
        mpole_exps)
    wrangler.communicate_mpoles(actx, mpole_exps)

    # (per-stage timing needs one stream: no overlap then)
    side = _side_stream(actx) if _OVERLAP_NEAR_FIELD and timing_data is None else None
    if side is None:
        timer.mark("eval_direct")
        potentials = wrangler.eval_direct(
            actx, trav.target_boxes, trav.neighbor_source_boxes_starts,
            trav.neighbor_source_boxes_lists, src_weight_vecs)

    timer.mark("multipole_to_local")
    local_exps = wrangler.multipole_to_local(
        actx, host("level_start_target_or_target_parent_box_nrs"),
        trav.target_or_target_parent_boxes, trav.from_sep_siblings_starts,
        trav.from_sep_siblings_lists, mpole_exps)

    if side is not None:
        # The near field (list 1) depends on nothing but the weights.  It
        # runs on a second stream once the M2L contraction has finished (run
        # beside it, the two compete for LDS and issue slots), so that it fills
        # the GPU while the latency-bound downward pass (lists 3 and 4, L2L,
        # L2P) runs; its potentials are joined at the end and summed in the
        # reference's order.
        import torch
        main = torch.cuda.current_stream(actx.device)
        side.wait_stream(main)
        with torch.cuda.stream(side):
            potentials = wrangler.eval_direct(
                actx, trav.target_boxes, trav.neighbor_source_boxes_starts,
                trav.neighbor_source_boxes_lists, src_weight_vecs)

    timer.mark("eval_multipoles")
    mpole_result = wrangler.eval_multipoles(
        actx, trav.target_boxes_sep_smaller_by_source_level,
        trav.from_sep_smaller_by_level, mpole_exps)
    # (a wrangler that knows its list-3 result is all zeros says so)
    add_mpole = not getattr(wrangler, "last_eval_multipoles_was_empty", False)

    close_smaller = None
    if getattr(trav, "from_sep_close_smaller_starts", None) is not None:
        close_smaller = wrangler.eval_direct(
            actx, trav.target_boxes, trav.from_sep_close_smaller_starts,
            trav.from_sep_close_smaller_lists, src_weight_vecs)

    timer.mark("form_locals")
    local_result = wrangler.form_locals(
        actx, host("level_start_target_or_target_parent_box_nrs"),
        trav.target_or_target_parent_boxes, trav.from_sep_bigger_starts,
        trav.from_sep_bigger_lists, src_weight_vecs)
    # (a wrangler that knows its list-4 result is all zeros says so)
    if not getattr(wrangler, "last_form_locals_was_empty", False):
        local_exps = local_exps + local_result

    close_bigger = None
    if getattr(trav, "from_sep_close_bigger_starts", None) is not None:
        close_bigger = wrangler.eval_direct(
            actx, trav.target_boxes, trav.from_sep_close_bigger_starts,
            trav.from_sep_close_bigger_lists, src_weight_vecs)

    timer.mark("refine_locals")
    local_exps = wrangler.refine_locals(
        actx, host("level_start_target_or_target_parent_box_nrs"),
        trav.target_or_target_parent_boxes, local_exps)

    timer.mark("eval_locals")
    local_result = wrangler.eval_locals(
        actx, host("level_start_target_box_nrs"), trav.target_boxes, local_exps)

    if side is not None:
        main.wait_stream(side)
        for pot in potentials:
            pot.record_stream(main)
        for w in src_weight_vecs:
            w.record_stream(side)
    if add_mpole:
        potentials = _add(potentials, mpole_result)
    if close_smaller is not None:
        potentials = _add(potentials, close_smaller)
    if close_bigger is not None:
        potentials = _add(potentials, close_bigger)
    potentials = _add(potentials, local_result)

    timer.mark("reorder_potentials")
    result = wrangler.reorder_potentials(potentials)
    result = wrangler.finalize_potentials(actx, result)
    timer.mark("end")
    if timing_data is not None:
        timing_data.update(timer.result())
    return result


_OVERLAP_NEAR_FIELD = True
_SIDE = {}


def _side_stream(actx):
    """A second (default-priority) stream on the context's device, or None
    off the GPU."""
    import torch
    dev = getattr(actx, "device", None)
    if dev is None or torch.device(dev).type != "cuda":
        return None
    key = str(torch.device(dev))
    if key not in _SIDE:
        _SIDE[key] = torch.cuda.Stream(torch.device(dev))
    return _SIDE[key]


def _add(a, b):
    out = np.empty(len(a), dtype=object)
    for i in range(len(a)):
        out[i] = a[i] + b[i]
    return out


class FMMGraph:
    """One FMM pass captured in a HIP graph (``torch.cuda.CUDAGraph``).

    The level loops of the pass are host loops that launch one kernel per
    level and stage (``sumpy/fmm.py:619-1056``); replaying the captured graph
    removes their launch overhead.  :meth:`replay` re-runs the pass on the
    current contents of the weight tensors given at capture (update them in
    place) and returns the potential tensors, which every replay overwrites.
    The pass is run twice eagerly before capture, so that every per-level plan
    and tile schedule is built (and cached) outside the captured region; all
    device memory the pass allocates comes from the graph's private pool.
    """

    def __init__(self, actx, wrangler, src_weight_vecs):
        import torch
        self.actx, self.wrangler = actx, wrangler
        self.weights = list(src_weight_vecs)
        side = torch.cuda.Stream(actx.device)
        side.wait_stream(torch.cuda.current_stream(actx.device))
        with torch.cuda.stream(side):
            for _ in range(2):
                drive_fmm(actx, wrangler, self.weights)
        torch.cuda.current_stream(actx.device).wait_stream(side)
        torch.cuda.synchronize(actx.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = drive_fmm(actx, wrangler, self.weights)

    def replay(self):
        self.graph.replay()
        return self.outputs


def capture_drive_fmm(actx, wrangler, src_weight_vecs):
    """Capture :func:`drive_fmm` of *wrangler* on *src_weight_vecs* into a
    :class:`FMMGraph`."""
    return FMMGraph(actx, wrangler, src_weight_vecs)
