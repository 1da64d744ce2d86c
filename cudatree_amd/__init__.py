"""MI355X-native random forest with CudaTree's interface.

Exports the names of cudatree/__init__.py:1-5 that belong to the fit/predict
path: RandomForestClassifier, convert_result, load_data and timer.
"""
import numpy as np

from .random_forest import RandomForestClassifier, timer

__version__ = "0.6"


def convert_result(tran_table, res):
    """Map compacted labels back through the label table
    (cudatree/random_forest.py:13-15)."""
    return np.asarray(tran_table)[np.asarray(res)]


def load_data(ds_name):
    """The scikit-learn-bundled datasets of the reference's tests
    (cudatree/datasource.py:9-23): digits, iris, diabetes."""
    from .datasets import load_sklearn

    return load_sklearn(ds_name)


__all__ = ["RandomForestClassifier", "convert_result", "load_data", "timer"]
