"""Argument checks of the host-side mirror of the reference's R interface
that run before any device allocation (no GPU needed): data groups
(R/data.R:127-140), the starting time (R/filter-support.R:94-104)."""
import numpy as np
import pytest

import dust2_amd as d2

DATA = {"time": np.array([4.0, 8.0, 12.0, 16.0]), "incidence": np.arange(1.0, 5.0)}


def test_number_of_data_groups():
    # tests/testthat/test-filter.R:341-347
    with pytest.raises(ValueError, match="Expected 'data' to have 2 groups, but it had 1"):
        d2.dust_filter_create(d2.sir(), 0, DATA, 10, n_groups=2, shared_data=False)
    # shared_data = TRUE takes one group of data for every parameter group
    d2.dust_filter_create(d2.sir(), 0, DATA, 10, n_groups=2, shared_data=True)
    two = {"time": DATA["time"], "incidence": np.column_stack([DATA["incidence"]] * 2)}
    with pytest.raises(ValueError, match="Expected 'data' to have 1 groups, but it had 2"):
        d2.dust_filter_create(d2.sir(), 0, two, 10, n_groups=2, shared_data=True)
    with pytest.raises(ValueError, match="Expected 'data' to have 3 groups, but it had 2"):
        d2.dust_unfilter_create(d2.sir(), 0, two, n_groups=3, shared_data=False)


def test_starting_time():
    # tests/testthat/test-filter.R:350-357
    with pytest.raises(ValueError,
                       match=r"'time_start' \(6\) is later than the first time in 'data' \(4\)"):
        d2.dust_filter_create(d2.sir(), 6, DATA, 10)
    with pytest.raises(ValueError, match=r"'time_start' \(4.5\) is later"):
        d2.dust_unfilter_create(d2.sirode(), 4.5, DATA)
    malaria = {"time": [4.0, 8.0], "tested": [2, 4], "positive": [1, 2]}
    with pytest.raises(ValueError, match=r"'time_start' \(5\) is later"):
        d2.dust_filter_create(d2.malaria(), 5, malaria, 10, dt=1)
