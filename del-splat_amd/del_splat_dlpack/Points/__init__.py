def load_ply(path_file: str):
    """Points/__init__.py:1-4 of the reference: (cap xyz [N,3], cap rgb [N,3])."""
    from ..del_splat_dlpack import points_load_ply

    return points_load_ply(path_file)
