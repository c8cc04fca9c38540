"""A dict-backed stand-in for the part of h5py that cardiax_amd.io uses
(File, create_dataset, groups by "a/b" paths, dataset slicing), for tests in
an interpreter without h5py. Datasets are NumPy arrays; datasets created
with a shape and no data start as zeros, as HDF5's default fill value does."""
import numpy as np


class Group:
    def __init__(self):
        self.items = {}

    def create_dataset(self, name, shape=None, dtype=None, data=None):
        head, _, rest = name.partition("/")
        if rest:
            g = self.items.setdefault(head, Group())
            return g.create_dataset(rest, shape=shape, dtype=dtype, data=data)
        if head in self.items:
            raise ValueError("dataset %s exists" % head)
        if data is not None:
            arr = np.array(data) if dtype is None else np.array(data, dtype=dtype)
        else:
            arr = np.zeros(shape, dtype=dtype or np.float32)
        self.items[head] = arr
        return arr

    def __getitem__(self, name):
        head, _, rest = name.partition("/")
        item = self.items[head]
        return item[rest] if rest else item

    def __contains__(self, name):
        try:
            self[name]
            return True
        except KeyError:
            return False

    def __iter__(self):
        return iter(self.items)

    def flat(self, prefix=""):
        """name -> array for every dataset below this group."""
        out = {}
        for k, v in self.items.items():
            if isinstance(v, Group):
                out.update(v.flat(prefix + k + "/"))
            else:
                out[prefix + k] = v
        return out


class File(Group):
    files = {}

    def __init__(self, path, mode="r"):
        super().__init__()
        if mode == "w":
            File.files[path] = self
        else:
            self.items = File.files[path].items
        self.closed = False

    def close(self):
        self.closed = True

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
