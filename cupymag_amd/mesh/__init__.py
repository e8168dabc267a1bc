"""mesh layer of cupymag_amd (mirrors cupymag/mesh)."""
