"""Host-side representation of a QuIDS state (reference: quids::iteration,
src/quids.hpp:149-335): SoA arrays of object bytes, padded offsets, sizes and
complex magnitudes."""
import numpy as np


def alignment_offset(size, align):
    """quids::get_alignment_offset, src/quids.hpp:93-102."""
    if align <= 1:
        return 0
    off = align - size % align
    return 0 if off == align else off


class HostState:
    """Objects, offsets (N+1), sizes and magnitudes of a state, on the host."""

    def __init__(self, objects=None, object_begin=None, object_size=None, magnitude=None, total_proba=1.0):
        self.objects = np.zeros(0, np.uint8) if objects is None else np.ascontiguousarray(objects, np.uint8)
        self.object_begin = (np.zeros(1, np.uint64) if object_begin is None
                             else np.ascontiguousarray(object_begin, np.uint64))
        self.object_size = (np.zeros(0, np.uint32) if object_size is None
                            else np.ascontiguousarray(object_size, np.uint32))
        self.magnitude = (np.zeros(0, np.complex128) if magnitude is None
                          else np.ascontiguousarray(magnitude, np.complex128))
        self.total_proba = float(total_proba)

    @property
    def num_object(self):
        return int(self.object_size.shape[0])

    @classmethod
    def from_objects(cls, objs, mags=None, align=8, total_proba=1.0):
        """Builds a state by iteration::append (src/quids.hpp:174-188) of each object."""
        n = len(objs)
        mags = [1.0] * n if mags is None else list(mags)
        sizes = np.array([len(o) for o in objs], np.uint32)
        begin = np.zeros(n + 1, np.uint64)
        off = 0
        for i, o in enumerate(objs):
            off += len(o) + alignment_offset(len(o), align)
            begin[i + 1] = off
        buf = np.zeros(off, np.uint8)
        for i, o in enumerate(objs):
            b = int(begin[i])
            buf[b:b + len(o)] = np.frombuffer(bytes(o), np.uint8)
        return cls(buf, begin, sizes, np.array(mags, np.complex128), total_proba)

    def append(self, obj, mag=1.0, align=8):
        objs = [self.get_object(i)[0] for i in range(self.num_object)] + [bytes(obj)]
        mags = list(self.magnitude) + [mag]
        st = HostState.from_objects(objs, mags, align, self.total_proba)
        self.__dict__.update(st.__dict__)

    def get_object(self, i):
        b = int(self.object_begin[i])
        return bytes(self.objects[b:b + int(self.object_size[i])]), complex(self.magnitude[i])

    def objects_list(self):
        return [self.get_object(i)[0] for i in range(self.num_object)]

    def mag_pairs(self):
        """Magnitudes as a contiguous float64 (N, 2) array."""
        return np.ascontiguousarray(self.magnitude.view(np.float64).reshape(-1, 2))

    def __repr__(self):
        return f"HostState(num_object={self.num_object}, bytes={int(self.object_begin[-1])}, total_proba={self.total_proba})"
