"""ID-indexed object registries (role of CAstraObjectManager,
include/astra/AstraObjectManager.h:55-119)."""
import itertools
import threading

from .log import AstraError

_ids = itertools.count(1)
_ids_lock = threading.Lock()


def next_id():
    with _ids_lock:
        return next(_ids)


class Registry:
    def __init__(self, what):
        self.what = what
        self.objects = {}
        self.lock = threading.Lock()

    def add(self, obj):
        i = next_id()
        with self.lock:
            self.objects[i] = obj
        return i

    def get(self, i):
        try:
            return self.objects[int(i)]
        except (KeyError, TypeError, ValueError):
            raise AstraError("%s not found: %r" % (self.what, i))

    def delete(self, ids):
        if not isinstance(ids, (list, tuple, set)):
            ids = [ids]
        with self.lock:
            for i in ids:
                obj = self.objects.pop(int(i), None)
                rel = getattr(obj, "release", None)
                if rel is not None:
                    rel()

    def clear(self):
        self.delete(list(self.objects.keys()))

    def info(self):
        """Print the object list as CAstraObjectManager::info does
        (src/AstraObjectManager.cpp:161-191): a header, one line per object
        (id, initialised flag, description), a closing rule."""
        rule = "-" * 41
        lines = ["id  init  description", rule]
        with self.lock:
            items = sorted(self.objects.items())
        for k, obj in items:
            desc = obj.description() if hasattr(obj, "description") else type(obj).__name__
            lines.append("%d \tv     %s" % (k, desc))
        lines.append(rule)
        print("\n".join(lines) + "\n")


data2d = Registry("Data object")
data3d = Registry("Data object")
projectors2d = Registry("Projector")
projectors3d = Registry("Projector")
algorithms = Registry("Algorithm")
matrices = Registry("Matrix")
