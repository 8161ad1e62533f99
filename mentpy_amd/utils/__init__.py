"""Data utilities and the expressivity sampler."""
from .expressivity import (expressivity_with_histogram, haar_probability_density_of_fidelities,
                           sample_probability_density_of_fidelities)
from .generate_data import (generate_haar_random_states, generate_random_dataset,
                            random_special_unitary, train_test_split)

__all__ = ["generate_haar_random_states", "generate_random_dataset", "train_test_split",
           "random_special_unitary", "expressivity_with_histogram",
           "haar_probability_density_of_fidelities", "sample_probability_density_of_fidelities"]
