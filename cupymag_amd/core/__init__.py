"""core layer of cupymag_amd (mirrors cupymag/core)."""
