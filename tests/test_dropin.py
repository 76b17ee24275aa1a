"""The reference's import paths resolve to this build after install_reference_aliases()."""


def test_reference_module_paths_resolve():
    import radiance_meshes_amd
    radiance_meshes_amd.install_reference_aliases()
    from delaunay_rasterization.internal.alphablend_tiled_slang_interp import AlphaBlendTiledRender
    from delaunay_rasterization.internal.render_grid import RenderGrid
    from delaunay_rasterization.internal.sort_by_keys import sort_by_keys_cub
    from delaunay_rasterization.internal.tile_shader_slang import VertexShader, vertex_and_tile_shader
    from delaunay_rasterization.internal.util import recombine_tensors, split_tensors
    from delaunay_rasterization.internal.render_err import render_err
    from radiance_meshes_amd import render_err as render_err_mod
    assert render_err is render_err_mod.render_err
    from radiance_meshes_amd import rasterizer
    assert AlphaBlendTiledRender is rasterizer.AlphaBlendTiledRender
    assert vertex_and_tile_shader is rasterizer.vertex_and_tile_shader
    assert VertexShader is rasterizer.VertexShader
    assert sort_by_keys_cub.sort_by_keys is rasterizer.sort_by_keys
    g = RenderGrid(840, 1280, 4, 4)
    assert (g.grid_height, g.grid_width) == (210, 320)
    import torch
    nt, t = split_tensors({"a": torch.zeros(1), "b": 3})
    d = recombine_tensors(nt, t)
    assert set(d) == {"a", "b"} and d["b"] == 3
