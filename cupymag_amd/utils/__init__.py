"""utils layer of cupymag_amd (mirrors cupymag/utils)."""
