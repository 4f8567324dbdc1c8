"""The host-side checks of test_host.py, collected again under the `gpu` mark
so that the GPU pass runs them against the library built for and shipped to
the GPU machine: the C-ABI exports, the .tns and SPLATT binary readers against
the oracle, split / sort, the synthetic generator and the CISS builder against
the oracle's reference layout and the hand-worked fixtures."""
import pytest

from test_host import (test_binary_reader_known_answer, test_binary_reader_zero_based_and_errors,  # noqa: F401
                       test_binary_roundtrip_widths, test_ciss_native_structure,
                       test_ciss_nnz_multiple_of_64_and_single_row, test_ciss_reference_layout_equals_oracle,
                       test_ciss_reference_layout_hand_worked, test_ciss_unsorted_input_is_sorted_internally,
                       test_cli_prep_writes_reference_file_set, test_coo_sort_is_stable_lexicographic,
                       test_coo_split_fractions_cover_and_seed, test_library_exports_every_header_symbol,
                       test_options_defaults_follow_reference, test_status_strings, test_synth_affine_range,
                       test_synth_is_seeded_distinct_and_exact, test_tns_reader_errors, test_tns_reader_matches_oracle,
                       test_tns_write_roundtrip, test_work_unit_defaults, test_ref_quirk_weights_equal_oracle)

pytestmark = pytest.mark.gpu
