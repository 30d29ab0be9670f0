"""The command line mirror of ticit's CLI (src/main.rs): argument rules
restated from its backend_is_choosable test (main.rs:221-273), and the
rate format of Rust's f64 Display. No device is used."""
import pytest

from ticit_amd import __main__ as cli


def test_arguments_follow_the_reference_rules():
    a = cli.parse_args(["circuit.ticit"])
    assert a.backend == "hip" and not a.count_only and a.shots == 1000 and a.seed == 1
    assert cli.parse_args(["circuit.ticit", "--count-only"]).count_only
    with pytest.raises(SystemExit):
        cli.parse_args(["circuit.ticit", "--exact-k", "2"])  # needs --records-out
    with pytest.raises(SystemExit):
        cli.parse_args(["circuit.ticit", "--records-out", "rows", "--count-only"])
    assert cli.parse_args(["circuit.ticit", "--records-out", "rows", "--exact-k", "0,1,2"]).exact_k == [0, 1, 2]
    a = cli.parse_args(["circuit.ticit", "--backend", "hip", "--records-out", "rows",
                        "--postselection-mask", "1"])
    assert a.postselection_mask == [1] and a.records_out == "rows"
    with pytest.raises(SystemExit):
        cli.parse_args(["circuit.ticit", "-n", "0"])  # shots range 1..


def test_unprovided_backends_and_exact_k_are_refused(capsys):
    assert cli.main(["circuit.ticit", "--backend", "cpu"]) == 2
    assert "use --backend hip" in capsys.readouterr().err
    assert cli.main(["circuit.ticit", "--records-out", "rows", "--exact-k", "1"]) == 2
    assert "not supported by the hip backend" in capsys.readouterr().err


@pytest.mark.parametrize("value,text", [(0.0, "0"), (1.0, "1"), (0.25, "0.25"), (1e-7, "0.0000001"),
                                        (1 / 3, "0.3333333333333333"), (float("nan"), "nan")])
def test_rates_print_as_rust_displays_them(value, text):
    assert cli.rate(value) == text
