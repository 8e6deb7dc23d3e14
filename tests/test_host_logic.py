

def test_parse_autotune_report():
    """the tm_autotune report format: trial lines, then the chosen shapes"""
    from timothy_amd import parse_autotune
    text = ("trial neighbors nbr_block=64 walk_window=0 ms=1.2500\n"
            "trial neighbors walk_f32=1 ms=1.0000\n"
            "trial tiles_sph tile_width=128 unroll=2 tiles_ms=0.1000 sph_ms=0.0800 ms=0.1800\n"
            "chosen precision=mixed nbr_block=256 walk_window=1024 walk_f32=1 sph_tiles=1 tile_width=128 "
            "sph_block=256 unroll=2 lanes=1\n")
    r = parse_autotune(text)
    assert [t["stage"] for t in r["trials"]] == ["neighbors", "neighbors", "tiles_sph"]
    assert r["trials"][2] == dict(stage="tiles_sph", tile_width=128, unroll=2, tiles_ms=0.1, sph_ms=0.08, ms=0.18)
    assert r["chosen"]["precision"] == "mixed" and r["chosen"]["tile_width"] == 128
    import pytest
    with pytest.raises(ValueError):
        parse_autotune("nonsense line\n")
