"""Argument and state errors through the C-ABI: every one comes back as an error code with a
message (AdscfError in Python) and leaves the context usable; none reaches a kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_context_arguments():
    import adscf_amd as amd

    for args in [(0, 8, 8, 1.0, 1.0, 1.0, 0.1), (8, 8, 8, -1.0, 1.0, 1.0, 0.1), (8, 8, 8, 1.0, 1.0, 1.0, 0.0),
                 (1024, 1024, 256, 1.0, 1.0, 1.0, 0.1)]:  # the last is above 2^27 points
        with pytest.raises(amd.AdscfError) as e:
            amd.Context(*args)
        assert e.value.code == amd.ERR_ARG
    with pytest.raises(amd.AdscfError):
        amd.Context(8, 8, 8, 1.0, 1.0, 1.0, 0.1, ntypes=5)


def test_loop_state_and_setup_errors():
    import adscf_amd as amd

    with amd.Context(8, 8, 8, 2.0, 2.0, 2.0, 0.05) as ctx:
        ctx.set_interactions(25.0, [[0.0, 12.0], [12.0, 0.0]], [1.0, 1.0])
        w = np.zeros((2, 8, 8, 8))
        with pytest.raises(amd.AdscfError) as e:
            ctx.scf_set_field(w)  # before scf_setup
        assert e.value.code == amd.ERR_STATE
        with pytest.raises(amd.AdscfError) as e:
            ctx.scf_iterate(5)
        assert e.value.code == amd.ERR_STATE
        for types, segs, m in [([1, 3], [10, 10], 10),    # monomer type out of range
                               ([1, 2], [10, -1], 10),    # negative segments
                               ([1, 2], [0, 0], 10),      # no segments
                               ([1, 2], [10, 10], 41),    # history deeper than 40
                               ([1] * 65, [1] * 65, 10)]:  # more than 64 blocks
            with pytest.raises(amd.AdscfError) as e:
                ctx.scf_setup(types, segs, mixing_dim=m)
            assert e.value.code == amd.ERR_ARG, (types, segs, m)
        # the context still works after the errors
        ctx.scf_setup([1, 2], [10, 10], mixing_dim=4)
        ctx.scf_set_field(np.random.default_rng(0).standard_normal((2, 8, 8, 8)))
        st = ctx.scf_iterate(3, tol=-1.0)
        assert st.iters == 3 and np.isfinite(st.accuracy)


def test_diagnostic_and_batch_errors():
    import adscf_amd as amd
    with amd.Context(8, 8, 8, 2.0, 2.0, 2.0, 0.05) as ctx:
        for n in (0, 41):
            with pytest.raises(amd.AdscfError) as e:
                ctx.device_lu_solve(np.eye(n), np.ones(n))
            assert e.value.code == amd.ERR_ARG
        with pytest.raises(amd.AdscfError):
            ctx.set_middle_axis(3)
    inst = [dict(grid=(20, 20, 20), box=(3.0, 3.0, 3.0), ds=0.05, blocks=[1, 2], segments=[10, 10], kappa=25.0,
                 chi=[[0.0, 12.0], [12.0, 0.0]], vol=[1.0, 1.0])]
    with pytest.raises(amd.AdscfError) as e:  # batched axis lengths are 8, 16, 24, 32, 48, 64, 128
        amd.Batch(inst, mixing_dim=4, lam_mix=0.2)
    assert e.value.code == amd.ERR_ARG
