"""YAML configuration loader (cupymag/core/config_loader.py:24-83): same
missing-file / bad-YAML errors, warning for missing sections, dotted
``get`` and output-directory resolution relative to the config file."""
import warnings
from pathlib import Path

import yaml


class ConfigLoader:
    def __init__(self, config_path):
        self.config_path = Path(config_path)
        self.config = self._load_config()

    def _load_config(self):
        if not self.config_path.exists():
            raise FileNotFoundError(
                f"Configuration file not found: {self.config_path}\n"
                f"Please copy an example config from the examples/ directory.")
        try:
            with open(self.config_path, "r") as f:
                config = yaml.safe_load(f)
        except yaml.YAMLError as e:
            raise ValueError(f"Invalid YAML syntax in {self.config_path}: {e}")
        self._validate_config(config)
        return config

    def _validate_config(self, config):
        for section in ("physics", "grid", "simulation", "output"):
            if section not in config:
                warnings.warn(f"Warning: Missing required section: {section}, necessary for micromagnetic "
                              f"simulations.", UserWarning)

    def get(self, key_path, default=None):
        value = self.config
        for key in key_path.split("."):
            if isinstance(value, dict) and key in value:
                value = value[key]
            else:
                return default
        return value

    def get_section(self, section):
        return self.config.get(section, {})

    def get_output_dir(self):
        output = self.get("output", {}) or {}
        directory = output.get("directory")
        if directory is None:
            return self.config_path.parent.resolve()
        p = Path(directory)
        if p.is_absolute():
            return p
        return (self.config_path.parent / p).resolve()
