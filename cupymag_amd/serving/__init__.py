"""Batched stepping of many simulations and the continuous-batching
hysteresis-sweep server."""
