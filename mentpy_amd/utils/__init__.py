"""Data utilities."""
from .generate_data import (generate_haar_random_states, generate_random_dataset,
                            random_special_unitary, train_test_split)

__all__ = ["generate_haar_random_states", "generate_random_dataset", "train_test_split",
           "random_special_unitary"]
