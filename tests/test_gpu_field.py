"""The PrimeField / CrtPair tests (src/field/prime_field.rs:91-442) collected a
second time under the `gpu` mark, so that the GPU run also executes them on
the machine the engine runs on. The field API is host code of the engine
library; the tests themselves live in test_field.py and run unmarked on the
CPU as well.
"""
import pytest

from test_field import (small_primes, test_admission_and_search_respect_the_prime_range,  # noqa: F401
                        test_admission_matches_trial_division, test_catalog_is_the_transform_catalog,
                        test_crt_pair_matches_wide_integers, test_documentation_examples,
                        test_field_arithmetic_matches_wide_integers, test_vector_operations_match_wide_integers)

pytestmark = pytest.mark.gpu
