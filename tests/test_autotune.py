"""Candidate tile shapes of the autotuner (clvisc_amd/autotune.py): every shape fits the
stage kernel's LDS and thread limits, the library's default tile is timed first, and the
0 chunk (the library's automatic chunk) is always among the chunk lengths."""
from clvisc_amd.autotune import candidates, _valid, TILE_THREADS, MAX_HALO_CELLS, MAX_FACES


def test_candidates_fit_the_kernel_limits():
    for dims in [(385, 385, 115), (161, 161, 41), (29, 31, 11), (301, 301, 1), (161, 161, 1), (97, 45, 7)]:
        nx, ny, nz = dims
        c = candidates(nx, ny, nz)
        assert c, dims
        for ty, tz, xc in c:
            hy, hz = (2 if ny > 1 else 0), (2 if nz > 1 else 0)
            assert ty * tz <= TILE_THREADS
            assert (ty + 2 * hy) * (tz + 2 * hz) <= MAX_HALO_CELLS
            assert not hy or (ty + 1) * tz <= MAX_FACES
            assert not hz or ty * (tz + 1) <= MAX_FACES
            assert 0 <= xc <= nx
        assert any(xc == 0 for _, _, xc in c)


def test_default_tile_comes_first():
    for dims, default in [((385, 385, 115), (10, 23)), ((301, 301, 1), (252, 1)), ((161, 161, 41), (11, 21))]:
        c = candidates(*dims, default=default)
        assert c[0][:2] == default
        assert _valid(dims[1], dims[2], *default)
        # each (tile, chunk) pair once
        assert len(set(c)) == len(c)
