"""`pyzen.vec` / `pyzen.mat`: the vector and matrix types of the reference's
Python module (src/binding/base_types.cpp:300-360 binds base::Vec and
base::Mat as Vec2f/Vec3f/Vec4f/Vec2i/Vec3i/..., Mat2f/Mat3f/Mat4f/..., with
create_* helpers and translate/rotate/scale).  Small numpy-backed classes:
enough for scripts that build scenes through pyzen.api (Scene.add_mesh takes
lists of Vec3f / Vec2f / Vec3i and a Mat4f, src/core/scene.h:144-160).
Matrices are row-major and act on column vectors (M * v)."""
import math

import numpy as np


class _Vec:
    N = 3
    DTYPE = np.float32

    def __init__(self, *args):
        if len(args) == 0:
            v = np.zeros(self.N)
        elif len(args) == 1 and np.ndim(args[0]) == 0:
            v = np.full(self.N, args[0])
        elif len(args) == 1:
            v = np.asarray(args[0]).reshape(-1)
        else:
            v = np.asarray(args)
        if v.size != self.N:
            raise ValueError("%s takes %d components" % (type(self).__name__, self.N))
        self.v = v.astype(self.DTYPE)

    def _wrap(self, a):
        out = type(self).__new__(type(self))
        out.v = np.asarray(a).astype(self.DTYPE)
        return out

    @staticmethod
    def _val(o):
        return o.v if isinstance(o, _Vec) else o

    def __add__(self, o):
        return self._wrap(self.v + self._val(o))

    def __sub__(self, o):
        return self._wrap(self.v - self._val(o))

    def __mul__(self, o):
        return self._wrap(self.v * self._val(o))

    __rmul__ = __mul__
    __radd__ = __add__

    def __truediv__(self, o):
        return self._wrap(self.v / self._val(o))

    def __rtruediv__(self, o):
        return self._wrap(self._val(o) / self.v)

    def __neg__(self):
        return self._wrap(-self.v)

    def __iadd__(self, o):
        self.v = (self.v + self._val(o)).astype(self.DTYPE)
        return self

    def __isub__(self, o):
        self.v = (self.v - self._val(o)).astype(self.DTYPE)
        return self

    def __imul__(self, o):
        self.v = (self.v * self._val(o)).astype(self.DTYPE)
        return self

    def __itruediv__(self, o):
        self.v = (self.v / self._val(o)).astype(self.DTYPE)
        return self

    def __eq__(self, o):
        return isinstance(o, _Vec) and np.array_equal(self.v, o.v)

    def __getitem__(self, i):
        if not 0 <= i < self.N:
            raise IndexError(i)
        return self.v[i].item()

    def __setitem__(self, i, x):
        if not 0 <= i < self.N:
            raise IndexError(i)
        self.v[i] = x

    def __len__(self):
        return self.N

    def __iter__(self):
        return iter(self.v.tolist())

    def __repr__(self):
        return "%s(%s)" % (type(self).__name__, ", ".join(str(x) for x in self.v.tolist()))

    def dot(self, o):
        return float(np.dot(self.v.astype(np.float64), o.v))

    def cross(self, o):
        return self._wrap(np.cross(self.v, o.v))

    def length_squared(self):
        return float(np.dot(self.v.astype(np.float64), self.v))

    def length(self):
        return math.sqrt(self.length_squared())

    def normalized(self):
        return self._wrap(self.v / self.length())

    def normalize(self):
        self.v = (self.v / self.length()).astype(self.DTYPE)
        return self

    def is_zero(self):
        return not np.any(self.v)

    def sum(self):
        return self.v.sum().item()

    def abs(self):
        return self._wrap(np.abs(self.v))

    def exp(self):
        return self._wrap(np.exp(self.v))

    def max_component(self):
        return self.v.max().item()

    def min_component(self):
        return self.v.min().item()

    def x(self):
        return self[0]

    def y(self):
        return self[1]

    def z(self):
        return self[2]

    def w(self):
        return self[3]

    r, g, b = x, y, z


def _vec_class(name, n, dtype):
    return type(name, (_Vec,), {"N": n, "DTYPE": dtype})


class _VecModule:
    """The `vec` submodule."""
    Vec2f = _vec_class("Vec2f", 2, np.float32)
    Vec2d = _vec_class("Vec2d", 2, np.float64)
    Vec2i = _vec_class("Vec2i", 2, np.int32)
    Vec3f = _vec_class("Vec3f", 3, np.float32)
    Vec3d = _vec_class("Vec3d", 3, np.float64)
    Vec3i = _vec_class("Vec3i", 3, np.int32)
    Vec4f = _vec_class("Vec4f", 4, np.float32)
    Vec4d = _vec_class("Vec4d", 4, np.float64)
    Vec4i = _vec_class("Vec4i", 4, np.int32)

    def __init__(self):
        for n in ("2", "3", "4"):
            for t in ("f", "d", "i"):
                cls = getattr(self, "Vec%s%s" % (n, t))
                setattr(self, "create_vec%s%s" % (n, t), (lambda c: lambda seq: c(list(seq)))(cls))

    @staticmethod
    def dot(a, b):
        return a.dot(b)

    @staticmethod
    def cross(a, b):
        return a.cross(b)

    @staticmethod
    def normalize(a):
        return a.normalized()

    @staticmethod
    def length(a):
        return a.length()


class _Mat:
    N = 4
    DTYPE = np.float32

    def __init__(self, *args):
        if len(args) == 0:
            m = np.eye(self.N)
        elif len(args) == 1:
            m = np.asarray(args[0])
        else:
            m = np.asarray(args)
        m = m.reshape(self.N, self.N)
        self.m = m.astype(self.DTYPE)

    def _wrap(self, a):
        out = type(self).__new__(type(self))
        out.m = np.asarray(a).astype(self.DTYPE)
        return out

    def __mul__(self, o):
        if isinstance(o, _Mat):
            return self._wrap(self.m.astype(np.float64) @ o.m)
        if isinstance(o, _Vec):
            if o.N == self.N:
                return o._wrap(self.m.astype(np.float64) @ o.v)
            if o.N == self.N - 1:   # a point: homogeneous coordinate 1
                h = self.m.astype(np.float64) @ np.append(o.v, 1.0)
                return o._wrap(h[:-1] / h[-1])
        return self._wrap(self.m * o)

    def __getitem__(self, ij):
        return self.m[ij]

    def __eq__(self, o):
        return isinstance(o, _Mat) and np.array_equal(self.m, o.m)

    def __repr__(self):
        return "%s(%s)" % (type(self).__name__, self.m.tolist())

    def transpose(self):
        return self._wrap(self.m.T)

    def inverse(self):
        return self._wrap(np.linalg.inv(self.m.astype(np.float64)))

    def as_array(self):
        return self.m.astype(np.float64)


def _mat_class(name, n, dtype):
    return type(name, (_Mat,), {"N": n, "DTYPE": dtype})


class _MatModule:
    """The `mat` submodule."""
    Mat2f = _mat_class("Mat2f", 2, np.float32)
    Mat2d = _mat_class("Mat2d", 2, np.float64)
    Mat3f = _mat_class("Mat3f", 3, np.float32)
    Mat3d = _mat_class("Mat3d", 3, np.float64)
    Mat4f = _mat_class("Mat4f", 4, np.float32)
    Mat4d = _mat_class("Mat4d", 4, np.float64)

    def __init__(self):
        for n in ("2", "3", "4"):
            for t in ("f", "d"):
                cls = getattr(self, "Mat%s%s" % (n, t))
                setattr(self, "create_mat%s%s" % (n, t), (lambda c: lambda seq: c(list(seq)))(cls))

    @staticmethod
    def identity(m):
        return m._wrap(np.eye(m.N))

    @staticmethod
    def transpose(m):
        return m.transpose()

    @staticmethod
    def inverse(m):
        return m.inverse()

    def translate3f(self, t):
        m = np.eye(4)
        m[:3, 3] = list(t)
        return self.Mat4f(m)

    def scale3f(self, s):
        return self.Mat4f(np.diag(list(s) + [1.0]))

    def rotate3f(self, axis_angle):
        """Rotation about (x, y, z) by the angle in the fourth component
        (radians), as Transform::rotate reads it (src/core/transform.h:40-45)."""
        a = list(axis_angle)
        ax = np.asarray(a[:3], np.float64)
        ax = ax / np.linalg.norm(ax)
        x, y, z = ax
        c, s = math.cos(a[3]), math.sin(a[3])
        C = 1.0 - c
        r = np.eye(4)
        r[:3, :3] = [[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]]
        return self.Mat4f(r)


vec = _VecModule()
mat = _MatModule()
