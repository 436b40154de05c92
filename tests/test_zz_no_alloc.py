"""flint_reactor_advance_noalloc (no counterpart in the reference; for callers that cannot allocate
device memory) against flint_reactor_advance through the C ABI: the same kernel on the same launch
shape, so every cell's result, counters and status are bit-identical, on two calls in a row (the
second with fewer cells, over the scratch the first one left).  A launch shape with more teams than
the static scratch holds is rejected, and the allocating entry point still runs it."""
import numpy as np
import pytest

from flint_amd import states as S

pytestmark = pytest.mark.gpu


def test_noalloc_matches_advance(mechs, cuda):
    import torch
    from flint_amd.native import FlintError, Mechanism, _p, check, lib
    md = mechs["TSR-Rich-31"]
    m = Mechanism.from_data(md)
    L = lib()
    for n in (3000, 700):
        st = S.radical_states(md.species, md.W, 900 + n, 0, n, ignition_half=True)
        X0 = torch.from_numpy(np.vstack([S.soa(st["Y"]), st["T"][None, :]])).to(cuda)
        p = torch.from_numpy(st["p"]).to(cuda)
        Xa, Xb = X0.clone(), X0.clone()
        sa = torch.zeros((8, n), dtype=torch.int32, device=cuda)
        sb = torch.full((8, n), -1, dtype=torch.int32, device=cuda)
        ra = torch.full((n,), -1, dtype=torch.int32, device=cuda)
        rb = torch.full((n,), -2, dtype=torch.int32, device=cuda)
        print(f"[gpu] reactor noalloc n={n}", flush=True)
        m.advance(Xa, p, mode=1, stats=sa, status=ra)
        torch.cuda.synchronize()
        check(L.flint_reactor_advance_noalloc(m._h, n, n, 1, _p(Xb), _p(p), 1e-4, 1e-6, 1e-12, 0, _p(sb), _p(rb),
                                              None))
        torch.cuda.synchronize()
        assert int((ra != 0).sum()) == 0
        assert torch.equal(Xa, Xb) and torch.equal(sa, sb) and torch.equal(ra, rb)
        assert not torch.equal(Xb, X0)
    # more teams than the 2048 Jacobian slots of the scratch: rejected, nothing written
    m.set_launch(8, 512)
    Xc = X0.clone()
    with pytest.raises(FlintError):
        check(L.flint_reactor_advance_noalloc(m._h, 700, 700, 1, _p(Xc), _p(p), 1e-4, 1e-6, 1e-12, 0, None, None,
                                              None))
    torch.cuda.synchronize()
    assert torch.equal(Xc, X0)
    m.advance(Xc, p, mode=1)
    torch.cuda.synchronize()
    assert torch.equal(Xc, Xa)
