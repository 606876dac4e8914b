"""MI355X-native mode-split step for the ZhenMode global ocean model."""
