"""Defect shape registry (cupymag/utils/defect_shapes.py:21-160): Sphere(R),
Ellipsoid((a,b,c)), RotatedEllipsoid(((a,b,c), rotation)).  Each shape
function maps (N,3) coordinates relative to the defect centre to a boolean
mask (strict inequality, as in the reference)."""
import numpy as np


class DefectShape:
    _shape_registry = {}

    @classmethod
    def register_shape(cls, name, shape_class):
        cls._shape_registry[name] = shape_class

    @classmethod
    def get_supported_shapes(cls):
        return list(cls._shape_registry.keys())


class SphereShape:
    def __init__(self, shape_params):
        self.R2 = float(shape_params) ** 2

    def create_shape_function(self):
        R2 = self.R2
        return lambda rel: (np.dot(rel, rel) < R2) if rel.ndim == 1 else (np.sum(rel ** 2, axis=1) < R2)


class EllipsoidShape:
    def __init__(self, shape_params):
        a, b, c = shape_params
        self.inv2 = np.array([1.0 / a ** 2, 1.0 / b ** 2, 1.0 / c ** 2])

    def create_shape_function(self):
        inv2 = self.inv2
        return lambda rel: (np.sum(inv2 * rel ** 2) < 1.0) if rel.ndim == 1 else \
            (np.sum(rel ** 2 * inv2[None, :], axis=1) < 1.0)


class RotatedEllipsoidShape:
    def __init__(self, shape_params):
        abc, rot = shape_params
        a, b, c = abc
        R = np.asarray(rot, dtype=float)
        self.Q = R @ np.diag([1.0 / a ** 2, 1.0 / b ** 2, 1.0 / c ** 2]) @ R.T

    def create_shape_function(self):
        Q = self.Q
        return lambda rel: (rel @ Q @ rel < 1.0) if rel.ndim == 1 else (np.sum((rel @ Q) * rel, axis=1) < 1.0)


DefectShape.register_shape("Sphere", SphereShape)
DefectShape.register_shape("Ellipsoid", EllipsoidShape)
DefectShape.register_shape("RotatedEllipsoid", RotatedEllipsoidShape)


def get_defect_shape_function(defect_shape, shape_params):
    # indexes the registry directly, so an unknown name raises KeyError like
    # the reference (defect_shapes.py:159)
    return DefectShape._shape_registry[defect_shape](shape_params).create_shape_function()
