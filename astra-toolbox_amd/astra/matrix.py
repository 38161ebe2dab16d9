"""Sparse-matrix objects (python/astra/matrix.py, python/astra/matrix_c.pyx:
CSparseMatrix, include/astra/SparseMatrix.h): CSR matrices with float32
values and 32-bit column indices, kept by id like the data objects.  The
reference's CPU projectors read them; here they are a host-side store.
"""
import numpy as np
import scipy.sparse as ss

from ._registry import matrices as _reg
from .log import AstraError

__all__ = ["create", "store", "get", "get_size", "delete", "clear", "info"]


class _Sparse:
    def __init__(self, shape, capacity):
        self.shape = (int(shape[0]), int(shape[1]))
        self.capacity = int(capacity)
        self.indptr = np.zeros(self.shape[0] + 1, np.int64)
        self.indices = np.zeros(self.capacity, np.uint32)
        self.data = np.zeros(self.capacity, np.float32)

    def put(self, m):
        """Copy a scipy sparse matrix in (csr_matrix_to_astra,
        python/astra/matrix_c.pyx:54-67): at most the capacity given at
        creation and at most as many rows."""
        if not ss.issparse(m):
            raise AstraError("Provided matrix object was not initialized")
        csr = m.tocsr()
        if csr.nnz > self.capacity or csr.shape[0] > self.shape[0]:
            raise AstraError("Matrix too large to store in this object")
        self.indptr[:len(csr.indptr)] = csr.indptr
        self.indices[:csr.nnz] = csr.indices
        self.data[:csr.nnz] = csr.data

    def csr(self):
        n = int(self.indptr[self.shape[0]])
        return ss.csr_matrix((self.data[:n].copy(), self.indices[:n].astype(np.int32),
                              self.indptr.copy()), shape=self.shape)

    def description(self):
        return "%d x %d, %d non-zeros" % (self.shape[0], self.shape[1], int(self.indptr[-1]))


def create(data):
    """A new matrix object holding `data` (a scipy sparse matrix); its id."""
    if not ss.issparse(data):
        raise AstraError("Provided matrix object was not initialized")
    obj = _Sparse(data.shape, data.nnz)
    obj.put(data)
    return _reg.add(obj)


def store(i, data):
    """Replace the matrix of object i (same rows, nnz within its capacity)."""
    _reg.get(i).put(data)


def get(i):
    """The matrix of object i as a scipy.sparse.csr_matrix (a copy)."""
    return _reg.get(i).csr()


def get_size(i):
    """(rows, columns) of matrix object i."""
    return _reg.get(i).shape


def delete(ids):
    _reg.delete(ids)


def clear():
    _reg.clear()


def info():
    _reg.info()
