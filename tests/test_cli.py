"""The command-line shim's error paths (no GPU needed: a deck error stops
before any device work, as the reference STOPs during setup)."""
import os

from alpineem_amd import cli


def test_cli_usage_without_arguments(capsys):
    assert cli.main([]) == 2
    assert "usage" in capsys.readouterr().err


def test_cli_reports_reference_stop_message(tmp_path, capsys, golden_dir):
    """A deck the reference rejects with STOP "ERROR: ..." prints that
    message and exits with status 1; the working directory is restored."""
    text = open(os.path.join(golden_dir, "decks", "dipole36.txt")).read().splitlines()
    text[0] = "36,36"                       # cell counts need three values
    deck = tmp_path / "inputs.txt"
    deck.write_text("\n".join(text) + "\n")
    cwd = os.getcwd()
    assert cli.main([str(deck)]) == 1
    assert os.getcwd() == cwd
    out = capsys.readouterr().out
    assert "Reading in the inputs file" in out


def test_cli_missing_deck(tmp_path, capsys):
    assert cli.main([str(tmp_path / "missing.txt")]) == 1
    assert capsys.readouterr().out.strip()
